"""MSGBSolver: MSGB forward and principal-symbol adjoint on one MI355X.

API of ``beamax.solvers.MSGBSolver`` (reference
``beamax/solvers/msgb_solvers/msgb_solver.py:356-1386``).  Host validation and
error messages follow the reference; every numerical step runs as a HIP
kernel through the C-ABI:

forward:  bh_mswpt_forward (0.5*a*half_mask) -> bh_topn -> bh_forward_params
          -> bh_forward_sum (mirrored beams, closed-form trajectories fused)
adjoint:  bh_adjoint_source -> bh_mswpt_forward (data tiling) -> bh_topn ->
          bh_adjoint_params (TR geometry, B^-1, propagation to t = 0) ->
          bh_static_sum (times c^-2)

``precision`` selects the summation arithmetic: "fp32_split" (FP64 phase,
FP32 cos/exp/partial sums, FP64 accumulation; returns float32) or "fp64"
(returns float64).  All other stages are FP64.
"""

from __future__ import annotations

import ctypes as C
import math
from numbers import Integral, Real
from typing import Optional, Tuple, Union

import numpy as np

from .. import _lib
from ..geometry import Domain, Sensor
from ..gb import gb_solvers
from ..transforms import MSWPT

_THR_STRATS = {"hard", "top_n", "percentile", "hard_reassign", "bao_energy", "perc_max_abs"}
_SUM_METHODS = {"all_real", "scan_real", "vmap_real", "all_complex", "scan_complex",
                "vmap_complex"}


def _validate_time_grid(ts, *, allow_singleton: bool = False) -> Optional[float]:
    """msgb_solver.py:41-65."""
    t = np.asarray(_np(ts), dtype=float)
    if t.ndim != 1 or t.size == 0:
        raise ValueError("ts must be a non-empty one-dimensional array.")
    if not np.all(np.isfinite(t)):
        raise ValueError("ts must contain only finite values.")
    if t.size == 1:
        if allow_singleton:
            return None
        raise ValueError("ts must be one-dimensional with at least two points.")
    dd = np.diff(t)
    if np.any(dd <= 0) or not np.allclose(dd, dd[0], rtol=1e-6, atol=0.0):
        raise ValueError("ts must be finite, strictly increasing, and uniformly spaced.")
    return float(dd[0])


def _np(x):
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def batch_params(params, batch_size: int, zero_padded_args=(4,)):
    """utils.batch_data (coeff_index.py:29-87) on torch tensors: pad to a batch
    multiple (zero for `zero_padded_args`, else repeat the last entry)."""
    import torch
    b = params[0].shape[0]
    nb = (b + batch_size - 1) // batch_size
    out = []
    for i, p in enumerate(params):
        pad = nb * batch_size - b
        if pad:
            if i in zero_padded_args:
                fill = torch.zeros((pad,) + tuple(p.shape[1:]), dtype=p.dtype, device=p.device)
            else:
                fill = p[-1:].expand((pad,) + tuple(p.shape[1:]))
            p = torch.cat([p, fill], dim=0)
        out.append(p.reshape((nb, batch_size) + tuple(p.shape[1:])))
    return tuple(out)


class ForwardBeams:
    """Device beam table (unique beams; the summation mirrors them to mode -1)."""

    def __init__(self, k: int, d: int, device):
        import torch
        self.count = int(k)
        self.d = d
        self.x0 = torch.empty((k, d), dtype=torch.float64, device=device)
        self.p0 = torch.empty((k, d), dtype=torch.float64, device=device)
        self.omega = torch.empty((k,), dtype=torch.float64, device=device)
        self.alpha = torch.empty((k, d), dtype=torch.complex128, device=device)
        self.a0 = torch.empty((k,), dtype=torch.complex128, device=device)

    def struct(self, offset: int = 0, count=None) -> "_lib.FwdBeams":
        """C view of rows [offset, offset + count) (all rows by default)."""
        n = self.count - offset if count is None else int(count)
        d = self.d
        return _lib.FwdBeams(n, self.x0.data_ptr() + 8 * d * offset,
                             self.p0.data_ptr() + 8 * d * offset,
                             self.omega.data_ptr() + 8 * offset,
                             self.alpha.data_ptr() + 16 * d * offset,
                             self.a0.data_ptr() + 16 * offset)


class StaticBeams:
    """Device table of beams frozen at one time level (adjoint beams at t = 0)."""

    def __init__(self, k: int, d: int, device):
        import torch
        self.count = int(k)
        self.d = d
        self.x = torch.empty((k, d), dtype=torch.float64, device=device)
        self.p = torch.empty((k, d), dtype=torch.float64, device=device)
        self.M = torch.empty((k, d, d), dtype=torch.complex128, device=device)
        self.A = torch.empty((k,), dtype=torch.complex128, device=device)
        self.omega = torch.empty((k,), dtype=torch.float64, device=device)

    def struct(self) -> "_lib.StaticBeams":
        return _lib.StaticBeams(self.count, self.x.data_ptr(), self.p.data_ptr(),
                                self.M.data_ptr(), self.A.data_ptr(), self.omega.data_ptr())

    @staticmethod
    def concat(tables):
        import torch
        t0 = tables[0]
        out = StaticBeams.__new__(StaticBeams)
        out.count = sum(t.count for t in tables)
        out.d = t0.d
        for f in ("x", "p", "M", "A", "omega"):
            setattr(out, f, torch.cat([getattr(t, f) for t in tables], dim=0).contiguous())
        return out


def topn_device(coeffs, k: int):
    """top_n on the GPU: (idx int64, values complex128), ascending (|c|, index)."""
    import torch
    n = coeffs.numel()
    k = min(int(k), n)
    if k <= 0:
        raise ValueError("top_n threshold must be a positive integer.")
    idx = torch.empty(k, dtype=torch.int64, device=coeffs.device)
    val = torch.empty(k, dtype=torch.complex128, device=coeffs.device)
    ws = _lib.workspace(_lib.load().bh_topn_workspace_size(n, k), coeffs.device)
    _lib.call("bh_topn", _lib.ptr(coeffs), n, k, _lib.ptr(idx), _lib.ptr(val), _lib.ptr(ws),
              ws.numel(), _lib.stream_handle())
    return idx, val


_THR_CODE = {"hard": 0, "percentile": 1, "hard_reassign": 2, "bao_energy": 3, "perc_max_abs": 4}


def threshold_device(coeffs, strategy: str, val, wpt: MSWPT = None):
    """threshold_coefficients (forward_solver_utils.py:181-230) on the GPU:
    (idx int64, values complex128).  top_n returns ascending (|c|, index);
    the other strategies return ascending indices, only the selected entries
    (the reference pads percentile/perc_max_abs to n with index -1 and value
    0, i.e. zero-amplitude beams).  The count is read back to the host, as
    the reference thresholds eagerly."""
    import torch
    if strategy == "top_n":
        return topn_device(coeffs, val)
    if strategy not in _THR_CODE:
        raise ValueError(f"Invalid thresholding strategy: {strategy}")
    if strategy == "bao_energy" and wpt is None:
        raise ValueError("bao_energy thresholding requires wpt.")
    n = coeffs.numel()
    dev = coeffs.device
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    vals = torch.empty(n, dtype=torch.complex128, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    ws = _lib.workspace(_lib.load().bh_threshold_workspace_size(n), dev)
    t = C.byref(wpt.tiling) if strategy == "bao_energy" else None
    _lib.call("bh_threshold", _lib.ptr(coeffs), n, _THR_CODE[strategy], float(val), t,
              _lib.ptr(idx), _lib.ptr(vals), _lib.ptr(cnt), _lib.ptr(ws), ws.numel(),
              _lib.stream_handle())
    k = int(cnt.item())
    return idx[:k].contiguous(), vals[:k].contiguous()


def detector_plane(positions, domain: Domain):
    """Normal axis, normal coordinate and inward sign of a planar detector
    (tr_solver_utils.py:245-250 find_constant_columns; 356-371 inward sign)."""
    pos = np.asarray(positions, dtype=float)
    const = np.all(pos == pos[0], axis=0)
    axes = np.flatnonzero(const)
    normal_axis = int(axes[0]) if axes.size else 0
    normal_value = float(pos[0, normal_axis]) if axes.size else 0.0
    dx = np.asarray(domain.dx, dtype=float)
    size = domain.grid_size
    gmax = size[normal_axis] - dx[normal_axis]
    close = lambda a, b: abs(a - b) <= 1e-8 + 1e-5 * abs(b)
    if close(normal_value, 0.0):
        inward = 1.0
    elif close(normal_value, gmax):
        inward = -1.0
    else:
        inward = float(np.sign(gmax / 2.0 - normal_value))
    if inward == 0.0:
        inward = 1.0
    return normal_axis, normal_value, inward


class MSGBSolver:
    """Multiscale Gaussian Beam solver (msgb_solver.py:356-534)."""

    def __init__(self, thr: Union[int, float], thr_strat: str, batch_size: int, input_type: str,
                 ode_solver, sum_method: str, tr_ode_solver=None, sharding=None,
                 ode_config=None, adjoint_relative_guard: float = 5e-2,
                 precision: str = "fp32_split", cull_log2: float = 40.0):
        if thr_strat not in _THR_STRATS:
            raise ValueError(f"thr_strat must be one of {', '.join(sorted(_THR_STRATS))}; "
                             f"got {thr_strat!r}.")
        if isinstance(thr, bool) or not isinstance(thr, Real):
            raise ValueError(f"thr must be a finite real scalar; got {thr!r}.")
        if not math.isfinite(float(thr)):
            raise ValueError(f"thr must be finite; got {thr!r}.")
        if thr_strat == "top_n":
            if not isinstance(thr, Integral) or int(thr) <= 0:
                raise ValueError("top_n threshold must be a positive integer.")
            thr = int(thr)
        elif thr_strat == "percentile":
            if not 0 <= float(thr) <= 100:
                raise ValueError("percentile threshold must lie in [0, 100].")
            thr = float(thr)
        elif thr_strat in {"hard_reassign", "perc_max_abs"}:
            if not 0 <= float(thr) <= 1:
                raise ValueError(f"{thr_strat} threshold must lie in [0, 1].")
            thr = float(thr)
        else:
            if float(thr) < 0:
                raise ValueError(f"{thr_strat} threshold must be non-negative.")
            thr = float(thr)
        if input_type not in {"spatial", "fourier"}:
            raise ValueError(f"input_type must be 'spatial' or 'fourier'; got {input_type!r}.")
        if batch_size <= 0:
            raise ValueError(f"batch_size must be positive; got {batch_size}.")
        if not 0.0 <= adjoint_relative_guard < 1.0:
            raise ValueError("adjoint_relative_guard must lie in [0, 1); got "
                             f"{adjoint_relative_guard}.")
        if sum_method not in _SUM_METHODS:
            raise ValueError(f"sum_method must be one of {', '.join(sorted(_SUM_METHODS))}; "
                             f"got {sum_method!r}.")
        if precision not in _lib.PRECISION:
            raise ValueError(f"precision must be 'fp32_split' or 'fp64'; got {precision!r}.")
        self.thr = thr
        self.thr_strat = thr_strat
        self.batch_size = batch_size
        self.input_type = input_type
        self.ode_solver = ode_solver
        self.tr_ode_solver = ode_solver if tr_ode_solver is None else tr_ode_solver
        self.sharding = sharding
        self.ode_config = ode_config if ode_config is not None else gb_solvers.SolverConfig.from_precision()
        self.adjoint_relative_guard = float(adjoint_relative_guard)
        self.use_real = "real" in sum_method
        self.aggregate_method = ("scan" if "scan" in sum_method else
                                 "vmap" if "vmap" in sum_method else "all")
        self.precision = precision
        # envelope cut: a (beam, 64-point group) pair is skipped when every
        # term is below 2^-cull_log2 of the beam's own 2|A|; 0 = dense sums
        self.cull_log2 = float(cull_log2)
        self._seg_cache = {}
        # optional device counter (uint64 tensor) of evaluated (beam, 64-point) pairs
        self.pair_counter = None

    # ------------------------------------------------------------------
    def _effective_top_n(self, wpt: MSWPT, *, half_frame: bool = False):
        """msgb_solver.py:536-549."""
        if self.thr_strat != "top_n":
            return self.thr
        cap = wpt.half_frame_capacity if half_frame else wpt.total_coeffs
        return min(int(self.thr), int(cap))

    def _select(self, coeffs, wpt: MSWPT, *, half_frame: bool = False):
        """threshold_coefficients with this solver's strategy and threshold."""
        return threshold_device(coeffs, self.thr_strat,
                                self._effective_top_n(wpt, half_frame=half_frame), wpt)

    def _check_supported(self, complex_sum_ok: bool = False):
        if not self.use_real and not complex_sum_ok:
            raise NotImplementedError("complex beam summation is built for the forward sum only; "
                                      "use *_real here.")
        if self.sharding is not None:
            raise NotImplementedError("multi-device sharding is not part of this single-GPU build")
        if self.ode_solver not in (gb_solvers.solve_hom_diag, gb_solvers.solve_ODE_base):
            raise NotImplementedError("the GPU forward runs solve_hom_diag (closed form, fused "
                                      "into the summation) or solve_ODE_base (adaptive RK in a "
                                      "grid medium)")

    @staticmethod
    def _c0(domain: Domain) -> float:
        """The speed the closed forms use: c at the origin
        (gb_solvers.py:356-422, c(zeros(d)))."""
        if domain.is_homogeneous:
            return float(domain.c)
        import torch
        sp = domain.speed_model
        return float(sp.evaluate(np.zeros((1, domain.ndim)))[0].reshape(-1)[0].item())

    @staticmethod
    def _sensor_positions(sensors):
        import torch
        if isinstance(sensors, Sensor):
            return np.asarray(sensors.positions, dtype=float)
        if isinstance(sensors, (np.ndarray, torch.Tensor)):
            return np.asarray(_np(sensors), dtype=float)
        return None

    # ------------------------------------------------------------------
    def prepare_forward_beams(self, p0, domain: Domain, wpt: MSWPT, dpdt=None):
        """compute_coefficients('pos_only') -> top_n -> compute_forward_parameters
        (msgb_solver.py:570-633); returns (ForwardBeams, idx, values)."""
        if dpdt is None:
            cpos = wpt.forward(p0, self.input_type, scale=0.5, half_mask=True)
        else:
            import torch
            a = wpt.forward(p0, self.input_type)
            b = wpt.forward(dpdt, self.input_type)
            cpos = torch.empty_like(a)
            _lib.call("bh_mswpt_combine_velocity", C.byref(wpt.tiling),
                      _lib.dbl_array(domain.grid_size), float(domain.c), _lib.ptr(a), _lib.ptr(b),
                      1, _lib.ptr(cpos), None, _lib.stream_handle())
        idx, val = self._select(cpos, wpt, half_frame=True)
        beams = ForwardBeams(idx.numel(), domain.ndim, wpt.device)
        gs = _lib.dbl_array(domain.grid_size)
        _lib.call("bh_forward_params", C.byref(wpt.tiling), gs, _lib.ptr(idx), _lib.ptr(val),
                  idx.numel(), beams.struct(), _lib.stream_handle())
        return beams, idx, val

    def _ts_device(self, ts, device):
        """The time grid on the device, cached by its contents: a repeated
        call makes no host-to-device copy (a pageable copy would wait for the
        stream's earlier work before the sum could be launched)."""
        import torch
        if isinstance(ts, torch.Tensor) and ts.device == torch.device(device) \
                and ts.dtype == torch.float64:
            return ts.contiguous()
        t = np.ascontiguousarray(_np(ts), dtype=np.float64)
        key = (t.shape, hash(t.tobytes()), str(device))
        cache = self.__dict__.setdefault("_ts_cache", {})
        hit = cache.get(key)
        if hit is None:
            hit = torch.as_tensor(t, device=device)
            cache[key] = hit
        return hit

    def _points_device(self, points, d, device):
        """Evaluation points on the device, cached by their contents (a
        repeated call neither copies them nor stalls a captured stream)."""
        import torch
        if isinstance(points, torch.Tensor) and points.device == torch.device(device) \
                and points.dtype == torch.float64:
            return points.reshape(-1, d).contiguous()
        pn = np.ascontiguousarray(_np(points), dtype=np.float64).reshape(-1, d)
        key = (pn.shape, hash(pn.tobytes()), str(device))
        cache = self.__dict__.setdefault("_pts_cache", {})
        hit = cache.get(key)
        if hit is None:
            hit = torch.as_tensor(pn, device=device)
            cache[key] = hit
        return hit

    def _segments(self, points, dx, device):
        from ..segments import Segments
        if hasattr(points, "data_ptr"):
            # a device tensor (the image grid of the adjoint): keyed by the
            # tensor itself and its in-place version, so a repeated call
            # neither copies it to the host nor hashes it (the copy would
            # stall the stream between the preparation and the sum)
            key = ("tensor", id(points), points.data_ptr(), points._version,
                   tuple(points.shape), tuple(dx), str(device))
            hit = self._seg_cache.get(key)
            if hit is None:
                pn = np.ascontiguousarray(_np(points), dtype=np.float64).reshape(-1, len(dx))
                hit = (Segments(pn, dx, device), points)   # holds the tensor: its id stays unique
                self._seg_cache[key] = hit
            return hit[0]
        key = (points.shape, hash(np.ascontiguousarray(points, dtype=np.float64).tobytes()),
               tuple(dx), str(device))
        seg = self._seg_cache.get(key)
        if seg is None:
            seg = Segments(points, dx, device)
            self._seg_cache[key] = seg
        return seg

    def sum_forward(self, beams: ForwardBeams, domain: Domain, sensor_positions, ts, out=None,
                    npos=None):
        """Beam sum at the sensors, (Nt, S): real (2 Re sum, ``*_real`` sum
        methods) or complex.  npos=None mirrors every beam (modes +1 and -1);
        otherwise beams [0, npos) have mode +1 and the rest mode -1."""
        import torch
        dev = beams.x0.device
        d = domain.ndim
        if npos is not None or not self.use_real:
            if any(domain.periodic):
                raise NotImplementedError("complex-input or complex-sum forward with periodic "
                                          "boundaries is not on the GPU path")
            pts = np.ascontiguousarray(sensor_positions, dtype=np.float64).reshape(-1, d)
            seg = self._segments(pts, domain.dx, dev)
            tsd = self._ts_device(ts, dev)
            f64 = self.precision == "fp64"
            if self.use_real:
                dt = torch.float64 if f64 else torch.float32
            else:
                dt = torch.complex128 if f64 else torch.complex64
            if out is None:
                out = torch.empty((tsd.numel(), seg.S), dtype=dt, device=dev)
            b = beams.struct()
            ss = seg.struct()
            ws = _lib.workspace(_lib.load().bh_forward_culled_workspace_size(
                beams.count, int(npos is None), tsd.numel(), seg.S), dev)
            _lib.call("bh_forward_sum_modes", C.byref(b), d, 0 if npos is None else int(npos),
                      int(npos is None), self._c0(domain), _lib.ptr(tsd), tsd.numel(), C.byref(ss),
                      self.cull_log2, _lib.PRECISION[self.precision], int(not self.use_real),
                      _lib.ptr(out), _lib.ptr(self.pair_counter), _lib.ptr(ws), ws.numel(),
                      _lib.stream_handle())
            return out
        if self.cull_log2 > 0 and not any(domain.periodic):
            pts = np.ascontiguousarray(sensor_positions, dtype=np.float64).reshape(-1, d)
            seg = self._segments(pts, domain.dx, dev)
            if seg.efficient:
                tsd = self._ts_device(ts, dev)
                dt = torch.float64 if self.precision == "fp64" else torch.float32
                if out is None:
                    out = torch.empty((tsd.numel(), seg.S), dtype=dt, device=dev)
                b = beams.struct()
                ss = seg.struct()
                ws = _lib.workspace(_lib.load().bh_forward_culled_workspace_size(
                    beams.count, 1, tsd.numel(), seg.S), dev)
                _lib.call("bh_forward_sum_culled", C.byref(b), d, 1, self._c0(domain), _lib.ptr(tsd),
                          tsd.numel(), C.byref(ss), self.cull_log2, _lib.PRECISION[self.precision],
                          _lib.ptr(out), _lib.ptr(self.pair_counter), _lib.ptr(ws), ws.numel(),
                          _lib.stream_handle())
                return out
        pos = self._points_device(sensor_positions, d, dev)
        tsd = self._ts_device(ts, dev)
        nt, S = tsd.numel(), pos.shape[0]
        dt = torch.float64 if self.precision == "fp64" else torch.float32
        if out is None:
            out = torch.empty((nt, S), dtype=dt, device=dev)
        per = _lib.int32_array([int(p) for p in domain.periodic])
        b = beams.struct()
        _lib.call("bh_forward_sum", C.byref(b), d, 1, self._c0(domain), _lib.ptr(tsd), nt,
                  _lib.ptr(pos), S, _lib.dbl_array(domain.grid_size), per,
                  _lib.PRECISION[self.precision], _lib.ptr(out), _lib.stream_handle())
        return out

    def forward(self, p0, domain: Domain, sensors, ts, wpt: MSWPT, *, dpdt=None):
        """Sensor data (Nt, *sensors.shape[:-1]) (msgb_solver.py:858-899)."""
        data, _ = self.forward_with_params(p0, domain, sensors, ts, wpt, dpdt=dpdt)
        return data

    def forward_with_params(self, p0, domain: Domain, sensors, ts, wpt: MSWPT, *, dpdt=None):
        """msgb_solver.py:901-982; params = (p0s, M0s, x0s, omegas, a0s, modes)."""
        import torch
        if tuple(p0.shape) != domain.N or (dpdt is not None and tuple(dpdt.shape) != domain.N):
            raise ValueError(f"p0 and dpdt must both have shape {domain.N}; got "
                             f"{tuple(p0.shape)} and "
                             f"{tuple(dpdt.shape) if dpdt is not None else tuple(p0.shape)}.")
        if wpt.dyadic_decomp.N != domain.N:
            raise ValueError("wpt grid shape must match domain.N.")
        _validate_time_grid(ts, allow_singleton=True)
        pos = self._sensor_positions(sensors)
        if pos is None:
            raise ValueError("Unsupported sensor type")
        cplx = lambda x: x.is_complex() if isinstance(x, torch.Tensor) else np.iscomplexobj(x)
        is_complex = cplx(p0) or (dpdt is not None and cplx(dpdt))
        if dpdt is not None:
            nz = (torch.count_nonzero(dpdt).item() if isinstance(dpdt, torch.Tensor)
                  else int(np.count_nonzero(np.asarray(dpdt))))
            if nz == 0:
                dpdt = None   # 0.5 * (a +- 1j * 0 / vg) == 0.5 * a exactly
        numerical = self.ode_solver is gb_solvers.solve_ODE_base
        if (not domain.is_homogeneous or numerical) and dpdt is not None:
            raise NotImplementedError("dpdt != 0 with the numerical ODE path (group velocity at "
                                      "the packet positions) is not on the GPU path")
        self._check_supported(complex_sum_ok=True)
        out_shape = tuple(np.asarray(_np(sensors.positions) if isinstance(sensors, Sensor)
                                     else pos).shape[:-1])
        if is_complex:
            beams, npos = self.prepare_forward_beams_complex(p0, domain, wpt, dpdt=dpdt)
            data = self.sum_forward(beams, domain, pos, ts, npos=npos)
            return (data.reshape((data.shape[0],) + out_shape),
                    self._params_tuple(beams, npos=npos))
        beams, _, _ = self.prepare_forward_beams(p0, domain, wpt, dpdt=dpdt)
        if numerical:
            if is_complex or not self.use_real:
                raise NotImplementedError("complex inputs or sums with solve_ODE_base are not on "
                                          "the GPU path")
            data = self.sum_forward_ode(beams, domain, pos, ts)
        else:
            data = self.sum_forward(beams, domain, pos, ts)
        data = data.reshape((data.shape[0],) + out_shape)
        return data, self._params_tuple(beams)

    def sum_forward_ode(self, beams: ForwardBeams, domain: Domain, sensor_positions, ts):
        """Mirrored real beam sum with the beam states integrated in the
        medium (solve_ODE_base, gb_solvers.py:941-1016): bh_ode_forward
        writes every beam's state at every time, bh_forward_sum_traj sums them
        with the same envelope cut as the closed-form path."""
        import torch
        dev = beams.x0.device
        d = domain.ndim
        sp = gb_solvers.speed_of(domain.c_fn) if not domain.is_homogeneous else None
        if sp is None:
            from ..utils.interp import GridSpeed
            sp = GridSpeed(np.full((1,) * d, float(domain.c)))
        tsd = self._ts_device(ts, dev)
        nt = tsd.numel()
        total = 2 * beams.count
        ns = 2 * d + 2 * d * d + 2
        if nt * total * ns * 8 > 96 * 2 ** 30:
            raise MemoryError(f"the trajectory table ({nt} x {total} beams) exceeds 96 GiB")
        traj = torch.empty((nt, total, ns), dtype=torch.float64, device=dev)
        status = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        cfg = self.ode_config
        sps = sp.struct(dev, lam=domain.lam)
        b = beams.struct()
        _lib.call("bh_ode_forward", C.byref(b), d, 1, C.byref(sps), _lib.ptr(tsd), nt,
                  float(cfg.rtol), float(cfg.atol), int(cfg.max_steps), _lib.ptr(traj),
                  _lib.ptr(status), _lib.stream_handle())
        if total and bool((status[:total] != 0).any()):
            raise RuntimeError("beam ODE integration failed: more than max_steps steps or a "
                               "non-finite state")
        pts = np.ascontiguousarray(sensor_positions, dtype=np.float64).reshape(-1, d)
        seg = self._segments(pts, domain.dx, dev)
        dt = torch.float64 if self.precision == "fp64" else torch.float32
        out = torch.empty((nt, seg.S), dtype=dt, device=dev)
        ws = _lib.workspace(_lib.load().bh_forward_sum_culled_workspace_size(nt, seg.S), dev)
        ss = seg.struct()
        _lib.call("bh_forward_sum_traj", C.byref(b), d, 1, _lib.ptr(traj), _lib.ptr(tsd), nt,
                  C.byref(ss), self.cull_log2, _lib.PRECISION[self.precision], _lib.ptr(out),
                  _lib.ptr(self.pair_counter), _lib.ptr(ws), ws.numel(), _lib.stream_handle())
        return out

    def prepare_forward_beams_complex(self, p0, domain: Domain, wpt: MSWPT, dpdt=None):
        """Complex-input beams (msgb_solver.py:635-689): cpos and cneg of
        compute_coefficients('both') thresholded separately (capacity
        total_coeffs each); returns (ForwardBeams [pos | neg], npos)."""
        import torch
        a = wpt.forward(p0, self.input_type)
        b = wpt.forward(dpdt, self.input_type) if dpdt is not None else None
        cpos = torch.empty_like(a)
        cneg = torch.empty_like(a)
        _lib.call("bh_mswpt_combine_velocity", C.byref(wpt.tiling),
                  _lib.dbl_array(domain.grid_size), float(domain.c), _lib.ptr(a), _lib.ptr(b), 0,
                  _lib.ptr(cpos), _lib.ptr(cneg), _lib.stream_handle())
        ip, vp = self._select(cpos, wpt)
        ineg, vneg = self._select(cneg, wpt)
        kp, kn = ip.numel(), ineg.numel()
        beams = ForwardBeams(kp + kn, domain.ndim, wpt.device)
        gs = _lib.dbl_array(domain.grid_size)
        for off, idx, val in ((0, ip, vp), (kp, ineg, vneg)):
            _lib.call("bh_forward_params", C.byref(wpt.tiling), gs, _lib.ptr(idx), _lib.ptr(val),
                      idx.numel(), beams.struct(off, idx.numel()), _lib.stream_handle())
        beams.idx = (ip, ineg)
        return beams, kp

    def solve_ivp(self, p0, dpdt, domain: Domain, wpt: MSWPT, sensors, ts):
        """Wave-equation IVP with initial velocity (msgb_solver.py:1113-1147)."""
        return self.forward(p0, domain, sensors, ts, wpt, dpdt=dpdt)

    def solve_ivp_with_params(self, p0, dpdt, domain: Domain, wpt: MSWPT, sensors, ts):
        return self.forward_with_params(p0, domain, sensors, ts, wpt, dpdt=dpdt)

    def _params_tuple(self, beams: ForwardBeams, npos=None):
        """Parameters in the reference's tuple order (batched for scan/vmap):
        mirrored (npos None) or [pos | neg] with modes +1 / -1."""
        import torch
        M0 = torch.diag_embed(beams.alpha)
        ones = torch.ones(beams.count, dtype=torch.float64, device=beams.x0.device)
        if npos is None:
            dup = lambda t: torch.cat([t, t], dim=0)
            params = (dup(beams.p0), dup(M0), dup(beams.x0), dup(beams.omega), dup(beams.a0),
                      torch.cat([ones, -ones]))
        else:
            modes = ones.clone()
            modes[int(npos):] = -1.0
            params = (beams.p0, M0, beams.x0, beams.omega, beams.a0, modes)
        if self.aggregate_method in ("scan", "vmap"):
            params = batch_params(params, self.batch_size, zero_padded_args=(4,))
        return params

    # ------------------------------------------------------------------
    def _validate_boundary_data(self, data, data_domain: Domain, sources, *, ts=None) -> None:
        """msgb_solver.py:806-856."""
        if tuple(data.shape) != data_domain.N:
            raise ValueError(f"Boundary data must have shape {data_domain.N}; got {tuple(data.shape)}.")
        if ts is not None:
            _validate_time_grid(ts)
            span = (data.shape[0] - 1) * data_domain.dx[0]
            t = np.asarray(_np(ts), dtype=float)
            if not np.isclose(float(t[-1] - t[0]), span, rtol=1e-6, atol=1e-12):
                raise ValueError(f"ts spans {float(t[-1] - t[0])}, but data_domain spans {span}.")
        if not isinstance(sources, Sensor):
            raise ValueError("sources must be a Sensor on a regular planar grid.")
        pos = np.asarray(sources.positions)
        stds = np.std(pos, axis=0)
        normal_axis = int(np.argmin(stds))
        tol = 1e-7 * max(1.0, float(np.max(sources.domain.grid_size)))
        if stds[normal_axis] > tol:
            raise ValueError("sources must lie on an axis-aligned planar surface.")
        counts = []
        for ax in [a for a in range(pos.shape[1]) if a != normal_axis]:
            vals = np.unique(pos[:, ax])
            if vals.size > 2 and not np.allclose(np.diff(vals), np.diff(vals)[0], rtol=1e-6, atol=tol):
                raise ValueError("sources must be uniformly spaced on the detector plane.")
            counts.append(int(vals.size))
        expected = tuple(counts) if counts else (1,)
        actual = tuple(data.shape[1:]) or (1,)
        if actual != expected:
            raise ValueError("Boundary data detector axes do not match the source grid: "
                             f"got {actual}, expected {expected}.")
        if math.prod(expected) != pos.shape[0]:
            raise ValueError("sources must form a complete Cartesian detector grid.")

    def adjoint_source(self, data, dt: float, c, window=None):
        """_form_adjoint_source (msgb_solver.py:68-120) on the GPU; returns (Nt, Ns...).

        ``c``: the sound speed at the detectors, a number or one value per
        detector column (device tensor)."""
        import torch
        dev = data.device
        x = data.to(torch.float64).contiguous()
        nt = x.shape[0]
        ns = x.numel() // nt
        out = torch.empty_like(x)
        if isinstance(c, torch.Tensor):
            csrc = c.to(device=dev, dtype=torch.float64).reshape(-1).contiguous()
            if csrc.numel() != ns:
                raise ValueError(f"c_at_sources has {csrc.numel()} detector samples, which does "
                                 f"not match {ns} detector columns")
        else:
            csrc = torch.full((ns,), float(c), dtype=torch.float64, device=dev)
        w = None
        w2d = 0
        if window is not None:
            w = torch.as_tensor(_np(window), device=dev).to(torch.float64).contiguous()
            w2d = int(w.ndim > 1)
        ws = _lib.workspace(_lib.load().bh_adjoint_source_workspace_size(nt, ns), dev)
        _lib.call("bh_adjoint_source", _lib.ptr(x), nt, ns, float(dt), _lib.ptr(w), w2d,
                  _lib.ptr(csrc), _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream_handle())
        return out

    def prepare_adjoint_beams(self, data, sources: Sensor, data_domain: Domain, data_wpt: MSWPT,
                              c: float, window=None, swap_axes: bool = False,
                              plane=None, medium=None) -> StaticBeams:
        """adjoint source -> data MSWPT -> top_n -> compute_adj_parameters ->
        solve_hom_TR to t = 0 (msgb_solver.py:1168-1233, 1347-1366).

        ``medium = (detector_speed, image_speed)`` (``GridSpeed`` models in
        the physical frame) selects the heterogeneous path: the source weight
        c(sources)^2, the parameters at c(x_T) and the beam ODEs from t_start
        to 0 in the image medium (msgb_solver.py:1352-1384); ``c`` is then
        unused."""
        if plane is None:
            plane = detector_plane(sources.positions, sources.domain)
        if medium is not None:
            c = medium[0].evaluate(self._physical(sources.positions, swap_axes),
                                   device=data_wpt.device)[0]
        src = self.adjoint_source(data, data_domain.dx[0], c, window)
        coeffs = data_wpt.forward(src, self.input_type)
        idx, val = self._select(coeffs, data_wpt)
        if medium is not None:
            beams = self._medium_beams(idx, val, data_wpt, data_domain, sources.domain.ndim, plane,
                                       swap_axes, False, medium)
        else:
            normal_axis, normal_value, inward = plane
            beams = StaticBeams(idx.numel(), sources.domain.ndim, data_wpt.device)
            _lib.call("bh_adjoint_params", C.byref(data_wpt.tiling),
                      _lib.dbl_array(data_domain.grid_size), _lib.ptr(idx), _lib.ptr(val),
                      idx.numel(), normal_axis, normal_value, inward, float(c),
                      self.adjoint_relative_guard, int(swap_axes), beams.struct(),
                      _lib.stream_handle())
        beams.idx, beams.val, beams.coeffs = idx, val, coeffs   # diagnostics
        return beams

    @staticmethod
    def _physical(positions, swap_axes: bool):
        """Detector positions in the physical frame (axes 0 and 1 swapped back)."""
        pos = np.asarray(_np(positions) if not isinstance(positions, np.ndarray) else positions,
                         dtype=np.float64)
        if swap_axes:
            pos = pos.copy()
            pos[..., [0, 1]] = pos[..., [1, 0]]
        return np.ascontiguousarray(pos)

    def _medium_beams(self, idx, val, data_wpt: MSWPT, data_domain: Domain, d: int, plane,
                      swap_axes: bool, time_reversal: bool, medium) -> StaticBeams:
        """Boundary beams in a heterogeneous medium: compute_TR_parameters /
        compute_adj_parameters at the local speed c(x_T) of the detector
        medium (tr_solver_utils.py:336-435, adjoint_solver_utils.py:185-220),
        then solve_ODE_batch_t from t_start to 0 through the image medium
        (gb_solvers.py:1017-1113)."""
        import torch
        dev = data_wpt.device
        normal_axis, normal_value, inward = plane
        k = idx.numel()
        beams = StaticBeams(k, d, dev)
        t_start = torch.empty(k, dtype=torch.float64, device=dev)
        mode = torch.empty(k, dtype=torch.float64, device=dev)
        det_speed, img_speed = medium
        sp_det = det_speed.struct(dev)
        _lib.call("bh_boundary_terminal", C.byref(data_wpt.tiling),
                  _lib.dbl_array(data_domain.grid_size), _lib.ptr(idx), _lib.ptr(val), k,
                  normal_axis, normal_value, inward, C.byref(sp_det),
                  self.adjoint_relative_guard, int(time_reversal), int(swap_axes),
                  beams.struct(), _lib.ptr(t_start), _lib.ptr(mode), _lib.stream_handle())
        cfg = self.ode_config
        status = torch.empty(max(k, 1), dtype=torch.int32, device=dev)
        sp_img = img_speed.struct(dev)
        b = beams.struct()
        _lib.call("bh_ode_static", C.byref(b), d, C.byref(sp_img), _lib.ptr(t_start), None,
                  _lib.ptr(mode), 0.0, float(cfg.rtol), float(cfg.atol), float(cfg.dt0 or 0.0),
                  int(cfg.max_steps), _lib.ptr(status), _lib.stream_handle())
        if k and bool((status[:k] != 0).any()):
            raise RuntimeError("time-reversal beam ODE integration failed: more than max_steps "
                               "steps or a non-finite state")
        beams.t_start, beams.mode = t_start, mode   # diagnostics
        return beams

    def _medium(self, sources: Sensor, domain: Domain):
        """(detector speed, image speed) for a heterogeneous problem, else None."""
        if domain.is_homogeneous and sources.domain.is_homogeneous:
            return None
        return gb_solvers.speed_of(sources.domain), gb_solvers.speed_of(domain)

    def sum_static(self, beams: StaticBeams, points, scale: float, out=None, dx=None):
        """out[P] = scale * sum_b field_b(points) (compute_TR_result at t = 0).

        With ``dx`` (the grid spacing of the points) and culling on, the
        points are segment-decomposed and the culled kernel is used."""
        import torch
        dev = beams.x.device
        d = beams.d
        if self.cull_log2 > 0 and dx is not None:
            if isinstance(points, torch.Tensor):
                seg = self._segments(points, dx, dev)
            else:
                seg = self._segments(np.ascontiguousarray(points, dtype=np.float64).reshape(-1, d),
                                     dx, dev)
            if seg.efficient:
                dt = torch.float64 if self.precision == "fp64" else torch.float32
                if out is None:
                    out = torch.empty(seg.S, dtype=dt, device=dev)
                ws = _lib.workspace(
                    _lib.load().bh_static_sum_culled_workspace_size(beams.count, d, seg.S), dev)
                b = beams.struct()
                ss = seg.struct()
                _lib.call("bh_static_sum_culled", C.byref(b), d, C.byref(ss), float(scale),
                          self.cull_log2, _lib.PRECISION[self.precision], _lib.ptr(out),
                          _lib.ptr(self.pair_counter), _lib.ptr(ws), ws.numel(),
                          _lib.stream_handle())
                return out
        pts = self._points_device(points, d, dev)
        P = pts.shape[0]
        dt = torch.float64 if self.precision == "fp64" else torch.float32
        if out is None:
            out = torch.empty(P, dtype=dt, device=dev)
        ws = _lib.workspace(_lib.load().bh_static_sum_workspace_size(beams.count, d), dev)
        b = beams.struct()
        _lib.call("bh_static_sum", C.byref(b), d, _lib.ptr(pts), P, float(scale),
                  _lib.PRECISION[self.precision], _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                  _lib.stream_handle())
        return out

    def prepare_tr_beams(self, data, sources: Sensor, data_domain: Domain, data_wpt: MSWPT,
                         c: float, swap_axes: bool = False, plane=None,
                         medium=None) -> StaticBeams:
        """compute_coefficients('both') (the cpos half, = 0.5 * MSWPT(data)) ->
        top_n -> compute_TR_parameters -> solve_hom_TR to t = 0
        (msgb_solver.py:691-746); ``medium`` as in ``prepare_adjoint_beams``."""
        coeffs = data_wpt.forward(data, self.input_type, scale=0.5)
        idx, val = self._select(coeffs, data_wpt)
        if plane is None:
            plane = detector_plane(sources.positions, sources.domain)
        if medium is not None:
            beams = self._medium_beams(idx, val, data_wpt, data_domain, sources.domain.ndim, plane,
                                       swap_axes, True, medium)
        else:
            normal_axis, normal_value, inward = plane
            beams = StaticBeams(idx.numel(), sources.domain.ndim, data_wpt.device)
            _lib.call("bh_tr_params", C.byref(data_wpt.tiling),
                      _lib.dbl_array(data_domain.grid_size), _lib.ptr(idx), _lib.ptr(val),
                      idx.numel(), normal_axis, normal_value, inward, float(c), int(swap_axes),
                      beams.struct(), _lib.stream_handle())
        beams.idx, beams.val, beams.coeffs = idx, val, coeffs
        return beams

    def time_reversal(self, data, domain: Domain, sensors, sources: Sensor, ts,
                      data_domain: Domain, data_wpt: MSWPT):
        """MSGB time-reversal reconstruction, times 2 (msgb_solver.py:984-1111)."""
        q, _ = self.time_reversal_with_params(data, domain, sensors, sources, ts, data_domain,
                                              data_wpt)
        return q

    def time_reversal_with_params(self, data, domain: Domain, sensors, sources: Sensor, ts,
                                  data_domain: Domain, data_wpt: MSWPT):
        import torch
        if any(domain.periodic):
            raise ValueError("The MSGB time reversal solver only supports free space boundary "
                             "conditions.")
        pos = self._sensor_positions(sensors)
        if pos is None:
            raise ValueError("Unsupported sensor type")
        self._validate_boundary_data(data, data_domain, sources, ts=ts)
        if self.sharding is not None:
            self._check_supported()
        medium = self._medium(sources, domain)
        if not isinstance(data, torch.Tensor):
            data = torch.as_tensor(np.asarray(data, dtype=np.float64), device=data_wpt.device)
        beams = self.prepare_tr_beams(data.to(data_wpt.device).to(torch.float64), sources,
                                      data_domain, data_wpt,
                                      None if medium else float(sources.domain.c),
                                      medium=medium)
        q = self.sum_static(beams, pos, 2.0, dx=domain.dx)
        shape = np.asarray(sensors.positions if isinstance(sensors, Sensor) else pos).shape[:-1]
        return q.reshape(tuple(shape)), beams

    def adjoint(self, data, domain: Domain, sensors, sources: Sensor, ts, data_domain: Domain,
                data_wpt: MSWPT, *, window=None):
        """Principal-symbol approximation of P* data (msgb_solver.py:1235-1297)."""
        q, _ = self.adjoint_with_params(data, domain, sensors, sources, ts, data_domain, data_wpt,
                                        window=window)
        return q

    def adjoint_with_params(self, data, domain: Domain, sensors, sources: Sensor, ts,
                            data_domain: Domain, data_wpt: MSWPT, *, window=None):
        """msgb_solver.py:1299-1386."""
        import torch
        if any(domain.periodic):
            raise ValueError("MSGBSolver.adjoint currently assumes non-periodic spatial "
                             "boundaries in the reconstruction domain.")
        pos = self._sensor_positions(sensors)
        if pos is None:
            raise ValueError("Unsupported sensor type for `sensors` in adjoint().")
        self._validate_boundary_data(data, data_domain, sources, ts=ts)
        if self.sharding is not None:
            self._check_supported()
        medium = self._medium(sources, domain)
        if not isinstance(data, torch.Tensor):
            data = torch.as_tensor(np.asarray(data, dtype=np.float64), device=data_wpt.device)
        c_src = None if medium else float(sources.domain.c)
        beams = self.prepare_adjoint_beams(data.to(data_wpt.device), sources, data_domain,
                                           data_wpt, c_src, window, medium=medium)
        if medium is None:
            q = self.sum_static(beams, pos, 1.0 / float(domain.c) ** 2, dx=domain.dx)
        else:
            # the c(x)^-2 image weight (msgb_solver.py:123-127, 1384)
            q = self.sum_static(beams, pos, 1.0, dx=domain.dx)
            pts = _np(pos) if isinstance(pos, torch.Tensor) else np.asarray(pos, dtype=np.float64)
            cimg = medium[1].evaluate(pts.reshape(-1, domain.ndim), device=q.device)[0]
            q = (q.to(torch.float64) / (cimg * cimg)).to(q.dtype)
        shape = np.asarray(sensors.positions if isinstance(sensors, Sensor) else pos).shape[:-1]
        return q.reshape(tuple(shape)), beams

"""Many independent MSGB forward solves in one summation launch, capturable
as one HIP graph.

The reference runs solves one at a time (one ``MSGBSolver.forward`` call per
grid).  ``BatchedForward`` takes a list of forward problems of different
grids, beam counts, time grids and sensors, preallocates every buffer once,
and issues per solve the (millisecond) MSWPT / top_n / parameter kernels and
activity windows, then a single ``bh_forward_batch_run`` launch that sums
all solves' beams
(``k_culled_batch``: one 1-D grid, solve s owning a contiguous block range).
Nothing in ``run`` allocates, copies from the host or synchronises, so
``capture`` records the whole batch into one ``torch.cuda.CUDAGraph`` and
``replay`` launches it again.  ``tune`` is the tile-shape autotuner: it
times the batched launch for each point-group tile (4, 8, 16 or 32 groups of
64 points per workgroup) and keeps the fastest, cached per batch signature.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .geometry import Domain
from .segments import Segments
from .solvers.msgb_solver import ForwardBeams, MSGBSolver, _np
from .transforms import MSWPT

_TUNED = {}


class ForwardProblem:
    """One forward solve: real p0 (device, float64), domain, MSWPT, sensor
    positions (S, d) and time grid."""

    def __init__(self, p0, domain: Domain, wpt: MSWPT, sensor_positions, ts, thr: int):
        self.p0, self.domain, self.wpt = p0, domain, wpt
        self.points = np.ascontiguousarray(sensor_positions, dtype=np.float64)
        self.ts = np.ascontiguousarray(_np(ts), dtype=np.float64)
        self.thr = int(thr)


class BatchedForward:
    def __init__(self, solver: MSGBSolver, problems, gpb=None):
        import torch
        if not problems:
            raise ValueError("empty batch")
        solver._check_supported()
        if solver.thr_strat != "top_n":
            raise NotImplementedError("the batched forward thresholds with top_n")
        self.solver = solver
        self.problems = problems
        ndims = {p.domain.ndim for p in problems}
        if len(ndims) != 1:
            raise ValueError("all solves of a batch must have the same dimension")
        self.ndim = ndims.pop()
        self.dev = problems[0].wpt.device
        lib = _lib.load()
        f64 = solver.precision == "fp64"
        self._keep = []
        self.items = []
        for p in problems:
            if any(p.domain.periodic) or not p.domain.is_homogeneous:
                raise NotImplementedError("batched solves: free-space homogeneous domains only")
            w = p.wpt
            k = min(p.thr, w.half_frame_capacity)
            it = {
                "cpos": torch.empty(w.total_coeffs, dtype=torch.complex128, device=self.dev),
                "mws": _lib.workspace(w._ws_bytes, self.dev),
                "idx": torch.empty(k, dtype=torch.int64, device=self.dev),
                "val": torch.empty(k, dtype=torch.complex128, device=self.dev),
                "tws": _lib.workspace(lib.bh_topn_workspace_size(w.total_coeffs, k), self.dev),
                "beams": ForwardBeams(k, self.ndim, self.dev),
                "ts": torch.as_tensor(p.ts, device=self.dev),
                "seg": Segments(p.points, p.domain.dx, self.dev),
                "gs": _lib.dbl_array(p.domain.grid_size),
                "k": k,
            }
            it["out"] = torch.empty((p.ts.size, it["seg"].S),
                                    dtype=torch.float64 if f64 else torch.float32, device=self.dev)
            # activity windows of the solve's entries (rebuilt each run;
            # zeroed so that the tables are valid indices even before)
            it["win"] = (torch.zeros(max(int(lib.bh_forward_windows_workspace_size(k, 1)), 8),
                                     dtype=torch.uint8, device=self.dev)
                         if solver.cull_log2 > 0 else None)
            self.items.append(it)
        self.nsolve = len(self.items)
        self._plan = _lib.workspace(lib.bh_forward_batch_plan_size(self.nsolve), self.dev)
        self.signature = tuple((tuple(p.domain.N), it["k"], p.ts.size, it["seg"].S)
                               for p, it in zip(problems, self.items)) + (solver.precision,
                                                                         solver.cull_log2)
        self.gpb = None
        self.nblocks = 0
        self.graph = None
        self._stale_windows = True
        self.set_tile(gpb if gpb is not None else _TUNED.get(self.signature, 8))

    def set_tile(self, gpb: int):
        """Build the device plan for point-group tile ``gpb`` (host copy + sync)."""
        arr = (_lib.BatchSolve * self.nsolve)()
        for i, (p, it) in enumerate(zip(self.problems, self.items)):
            arr[i].beams = it["beams"].struct()
            arr[i].c = float(p.domain.c)
            arr[i].ts = it["ts"].data_ptr()
            arr[i].nt = p.ts.size
            arr[i].seg = it["seg"].struct()
            arr[i].out = it["out"].data_ptr()
            arr[i].win = it["win"].data_ptr() if it["win"] is not None else None
        self._solves = arr
        nb = C.c_int64(0)
        _lib.call("bh_forward_batch_plan", arr, self.nsolve, self.ndim, self.solver.cull_log2,
                  int(gpb), _lib.ptr(self._plan), self._plan.numel(), C.byref(nb),
                  _lib.stream_handle())
        self.gpb = int(gpb)
        self.nblocks = int(nb.value)
        self.graph = None

    def prepare(self):
        """Per solve: 0.5 * half-masked MSWPT, top_n, beam parameters."""
        self._stale_windows = True
        s = _lib.stream_handle()
        for p, it in zip(self.problems, self.items):
            w = p.wpt
            w.forward(p.p0, "spatial", scale=0.5, half_mask=True, out=it["cpos"], ws=it["mws"])
            _lib.call("bh_topn", _lib.ptr(it["cpos"]), w.total_coeffs, it["k"],
                      _lib.ptr(it["idx"]), _lib.ptr(it["val"]), _lib.ptr(it["tws"]),
                      it["tws"].numel(), s)
            _lib.call("bh_forward_params", C.byref(w.tiling), it["gs"], _lib.ptr(it["idx"]),
                      _lib.ptr(it["val"]), it["k"], it["beams"].struct(), s)

    def windows(self):
        """Every solve's activity windows (the summation then skips the rows
        on which an entry reaches no sensor)."""
        _lib.call("bh_forward_batch_windows", self._solves, self.nsolve, self.ndim,
                  self.solver.cull_log2, _lib.stream_handle())
        self._stale_windows = False

    def sum(self):
        """The one batched summation launch (after ``prepare`` and ``windows``)."""
        if self._stale_windows and self.solver.cull_log2 > 0:
            raise RuntimeError("BatchedForward.sum: the beams changed since the last "
                               "windows() (run prepare(), windows(), sum(), or run())")
        _lib.call("bh_forward_batch_run", _lib.ptr(self._plan), self.nsolve, self.nblocks,
                  self.ndim, _lib.PRECISION[self.solver.precision], self.gpb, None,
                  _lib.stream_handle())

    def run(self):
        self.prepare()
        self.windows()
        self.sum()
        return self.outputs

    @property
    def outputs(self):
        return [it["out"] for it in self.items]

    def dense_evals(self) -> int:
        return int(sum(2 * it["k"] * p.ts.size * it["seg"].S
                       for p, it in zip(self.problems, self.items)))

    def capture(self):
        """Record prepare + windows + sum into one HIP graph (torch.cuda.CUDAGraph)."""
        import torch
        self.run()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.outputs

    def tune(self, tiles=None, reps: int = 2):
        """Time the batched launch per tile shape; keep and cache the fastest."""
        import torch
        self.prepare()
        self.windows()
        best = None
        times = {}
        if tiles is None:
            # the FP32-split kernels take 8, 16 or 32 groups per block
            tiles = (4, 8, 16, 32) if self.solver.precision == "fp64" else (8, 16, 32)
        for gpb in tiles:
            self.set_tile(gpb)
            self.sum()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                self.sum()
            e1.record()
            torch.cuda.synchronize()
            times[gpb] = e0.elapsed_time(e1) / reps
            if best is None or times[gpb] < times[best]:
                best = gpb
        self.set_tile(best)
        _TUNED[self.signature] = best
        return best, times

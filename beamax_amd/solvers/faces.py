"""Adjoint of 2-D four-edge sensing as a sum of single-face adjoints.

The reference adjoint (``MSGBSolver.adjoint``, msgb_solver.py:1299-1386)
accepts one planar detector face and converts the spacetime Hessian treating
spatial axis 0 as the detector normal (tr_solver_utils.py:143-149).  A
four-edge detector is therefore split into its four faces:

* each face's columns are taken from the (Nt, S) record, corner sensors (which
  belong to two faces) with weight 1/2;
* the record is Fourier-cropped in time to ``nt_data`` samples exactly as
  ``prepare_data_domain_for_msgb`` does
  (examples/reconstruction/2d_time_reversal_and_adjoint.py:111-165);
* faces whose normal is axis 1 are solved in the frame with axes 0 and 1
  swapped and their beams are swapped back; in a heterogeneous medium the
  speed is read at the physical position of each detector point, and the
  beams are integrated to t = 0 in the physical frame;
* the beams of all four faces are summed over the image in one launch
  (then weighted by c(x)^-2).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from ..decomposition import DyadicDecomposition
from ..geometry import Domain, Sensor
from ..transforms import MSWPT
from .msgb_solver import MSGBSolver, StaticBeams, detector_plane


def four_edge_mask(N):
    """Binary mask of the four edges of an (n0, n1) grid."""
    m = np.zeros(N)
    m[0, :] = 1
    m[-1, :] = 1
    m[:, 0] = 1
    m[:, -1] = 1
    return m


def four_edge_faces(N):
    """Column lists, weights and plane geometry of the faces of a four-edge
    sensor array in row-major mask order (geometry.py:575-590)."""
    n0, n1 = N
    row_last = n1 + 2 * (n0 - 2)
    left = [0] + [n1 + 2 * (r - 1) for r in range(1, n0 - 1)] + [row_last]
    right = [n1 - 1] + [n1 + 2 * (r - 1) + 1 for r in range(1, n0 - 1)] + [row_last + n1 - 1]
    faces = [
        dict(name="row0", cols=list(range(n1)), swap=False, normal=0.0),
        dict(name="rowN", cols=[row_last + j for j in range(n1)], swap=False, normal=None),
        dict(name="col0", cols=left, swap=True, normal=0.0),
        dict(name="colN", cols=right, swap=True, normal=None),
    ]
    for f in faces:
        w = np.ones(len(f["cols"]))
        w[0] = 0.5
        w[-1] = 0.5
        f["weights"] = w
    return faces


class FourEdgeAdjoint:
    """P* for 2-D four-edge sensing (see module doc); data (Nt, S) -> image."""

    def __init__(self, solver: MSGBSolver, domain: Domain, ts, nt_data: int,
                 num_levels: int = 2, num_boxes_levels=(4, 8), device=None):
        import torch
        if domain.ndim != 2:
            raise ValueError("four-edge adjoint is defined for 2-D domains")
        self.solver = solver
        self.domain = domain
        ts = np.asarray(ts, dtype=float)
        self.nt = ts.size
        self.nt_data = int(nt_data)
        span = float(ts[-1] - ts[0])
        self.dt_data = span / (self.nt_data - 1)
        hom = domain.is_homogeneous
        self.c = float(domain.c) if hom else None
        self.medium = None if hom else (domain.speed_model, domain.speed_model)
        c_face = self.c if hom else 1.0   # the face domains' speed is unused when heterogeneous
        n0, n1 = domain.N
        self.faces = four_edge_faces(domain.N)
        self.S = 2 * n1 + 2 * (n0 - 2)
        wpts = {}
        for f in self.faces:
            rot = f["swap"]
            nf = len(f["cols"])
            N_rot = (n1, n0) if rot else (n0, n1)
            dx_rot = (domain.dx[1], domain.dx[0]) if rot else domain.dx
            dom_rot = Domain(N=N_rot, dx=dx_rot, periodic=(False, False), c=c_face, cfl=domain.cfl)
            tan_dx = dx_rot[1]
            normal_val = 0.0 if f["normal"] == 0.0 else (N_rot[0] - 1) * dx_rot[0]
            pos = np.stack([np.full(nf, normal_val), np.arange(nf) * tan_dx], axis=1)
            f["sources"] = Sensor(dom_rot, positions=pos)
            f["plane"] = detector_plane(pos, dom_rot)
            nmin = min(self.nt_data, nf)
            aspect = (self.nt_data // nmin, nf // nmin)
            f["data_domain"] = Domain(N=(self.nt_data, nf), dx=(self.dt_data, tan_dx),
                                      periodic=(False, False), c=c_face, cfl=domain.cfl)
            key = (self.nt_data, nf, aspect)
            if key not in wpts:
                dd = DyadicDecomposition(num_levels=num_levels, N=(self.nt_data, nf),
                                         num_boxes_levels=num_boxes_levels, box_aspect_ratio=aspect)
                wpts[key] = MSWPT(dd, redundancy=2, windowing="rectangular_mirror", device=device)
            f["wpt"] = wpts[key]
            dev = f["wpt"].device
            f["cols_dev"] = torch.as_tensor(np.asarray(f["cols"], dtype=np.int64), device=dev)
            f["w_dev"] = torch.as_tensor(f["weights"], dtype=torch.float64, device=dev)
        self.device = self.faces[0]["wpt"].device
        self.points = torch.as_tensor(np.ascontiguousarray(domain.grid.reshape(-1, 2)),
                                      device=self.device)
        if hom:
            self.weight = None
        else:
            cimg = self.medium[1].evaluate(self.points, device=self.device)[0]
            self.weight = 1.0 / (cimg * cimg)

    def crop(self, data, face):
        import torch
        x = data.to(torch.float64).contiguous()
        nf = len(face["cols"])
        out = torch.empty((self.nt_data, nf), dtype=torch.float64, device=x.device)
        ws = _lib.workspace(
            _lib.load().bh_fourier_crop_time_workspace_size(self.nt, self.nt_data, nf), x.device)
        _lib.call("bh_fourier_crop_time", _lib.ptr(x), self.nt, x.shape[1],
                  _lib.ptr(face["cols_dev"]), _lib.ptr(face["w_dev"]), nf, self.nt_data,
                  _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream_handle())
        return out

    def beams(self, data) -> StaticBeams:
        """The four faces' boundary beams, each face prepared on a stream of
        its own (their small crop/MSWPT/top_n/parameter kernels overlap);
        the caller's stream waits for all four before the concatenation."""
        import torch
        main = torch.cuda.current_stream(self.device)
        data = data.to(torch.float64).contiguous()   # once, on the caller's stream
        if getattr(self, "_streams", None) is None:
            self._streams = [torch.cuda.Stream(device=self.device) for _ in self.faces]
        tabs = []
        for f, st in zip(self.faces, self._streams):
            st.wait_stream(main)
            with torch.cuda.stream(st):
                face_data = self.crop(data, f)
                tabs.append(self.solver.prepare_adjoint_beams(
                    face_data, f["sources"], f["data_domain"], f["wpt"], self.c,
                    swap_axes=f["swap"], plane=f["plane"], medium=self.medium))
        for st in self._streams:
            main.wait_stream(st)
        self.face_beams = tabs   # per-face tables (selection diagnostics)
        return StaticBeams.concat(tabs)

    def __call__(self, data):
        import torch
        b = self.beams(data)
        if self.weight is None:
            q = self.solver.sum_static(b, self.points, 1.0 / self.c ** 2, dx=self.domain.dx)
        else:
            q = self.solver.sum_static(b, self.points, 1.0, dx=self.domain.dx)
            q = (q.to(torch.float64) * self.weight).to(q.dtype)
        return q.reshape(self.domain.N)

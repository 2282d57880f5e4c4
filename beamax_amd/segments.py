"""Decomposition of evaluation points into straight, equally spaced runs.

The culled summation kernels (``bh_forward_sum_culled``,
``bh_static_sum_culled``) test each beam against each run in closed form, so
the points (sensors or image pixels) are grouped here once per geometry:
points on the grid are split greedily into maximal runs along the last axis
(length >= ``min_run``), then along the earlier axes, and whatever is left
into runs of any length.  Each run is packed into groups of 64 points (one
per lane); padding slots repeat a valid point and write nowhere.  Every run
holds an even number of groups (an all-padding group completes an odd
count), so the kernels can pair adjacent groups of one run.
"""

from __future__ import annotations

import numpy as np

from . import _lib


class Segments:
    """Device tables of a segment decomposition (see ``bh_segments``)."""

    def __init__(self, positions, dx, device, min_run: int = 8):
        import torch
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64))
        if pos.ndim != 2:
            raise ValueError("positions must be (S, d)")
        S, d = pos.shape
        dx = np.asarray(dx, dtype=np.float64)
        idx = np.rint(pos / dx).astype(np.int64)
        self.on_grid = bool(np.all(np.abs(pos - idx * dx) <= 1e-6 * dx))
        runs = _runs(idx, min_run) if self.on_grid else [(np.array([i]), 0) for i in range(S)]
        seg_p0, seg_v, seg_n, seg_g0, grp_seg = [], [], [], [], []
        pts, col = [], []
        g = 0
        for k, (members, axis) in enumerate(runs):
            n = len(members)
            v = np.zeros(d)
            if n > 1:
                v[axis] = dx[axis]
            seg_p0.append(pos[members[0]])
            seg_v.append(v)
            seg_n.append(n)
            seg_g0.append(g)
            ng = (n + 63) // 64
            ng += ng % 2
            for q in range(ng):
                m = members[q * 64:(q + 1) * 64]
                if len(m):
                    slot_pts = np.repeat(pos[m[-1]][None], 64, axis=0)
                    slot_pts[:len(m)] = pos[m]
                else:   # padding group: the run continued past its end
                    slot_pts = pos[members[0]][None] + (q * 64 + np.arange(64))[:, None] * v
                slot_col = np.full(64, -1, dtype=np.int32)
                slot_col[:len(m)] = m
                pts.append(slot_pts)
                col.append(slot_col)
                grp_seg.append(k)
            g += ng
        self.nseg = len(runs)
        self.ngroups = g
        self.S = S
        self.d = d
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=dt)), device=device)
        self.seg_p0 = t(seg_p0, np.float64)
        self.seg_v = t(seg_v, np.float64)
        self.seg_n = t(seg_n, np.int32)
        self.seg_g0 = t(seg_g0, np.int32)
        self.grp_seg = t(grp_seg, np.int32)
        self.pts = t(np.concatenate(pts, axis=0), np.float64)
        self.col = t(np.concatenate(col, axis=0), np.int32)
        self.mean_run = S / max(1, self.nseg)

    @property
    def efficient(self) -> bool:
        """True when the groups are mostly full (culled kernels pay off)."""
        return self.on_grid and self.S >= 0.5 * 64 * self.ngroups

    def struct(self) -> "_lib.Segs":
        return _lib.Segs(self.nseg, self.ngroups, self.S, self.seg_p0.data_ptr(),
                         self.seg_v.data_ptr(), self.seg_n.data_ptr(), self.seg_g0.data_ptr(),
                         self.grp_seg.data_ptr(), self.pts.data_ptr(), self.col.data_ptr())


def _runs(idx, min_run):
    """Greedy maximal runs of consecutive grid indices; returns [(members, axis)]."""
    S, d = idx.shape
    left = np.ones(S, dtype=bool)
    out = []
    passes = [(a, min_run) for a in range(d - 1, -1, -1)] + [(d - 1, 1)]
    for axis, lmin in passes:
        cand = np.flatnonzero(left)
        if cand.size == 0:
            break
        other = [a for a in range(d) if a != axis]
        keys = [idx[cand, axis]] + [idx[cand, a] for a in reversed(other)]
        order = cand[np.lexsort(keys)]
        if order.size == 0:
            continue
        same = np.ones(order.size - 1, dtype=bool)
        for a in other:
            same &= idx[order[1:], a] == idx[order[:-1], a]
        step = (idx[order[1:], axis] - idx[order[:-1], axis]) == 1
        brk = np.flatnonzero(~(same & step)) + 1
        starts = np.concatenate([[0], brk])
        ends = np.concatenate([brk, [order.size]])
        for s0, e0 in zip(starts, ends):
            if e0 - s0 >= lmin:
                members = order[s0:e0]
                out.append((members, axis))
                left[members] = False
    assert not left.any()
    return out

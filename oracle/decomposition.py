"""Oracle: dyadic Fourier tiling (test infrastructure only).

Restates ``beamax/decomposition.py`` (reference) in plain NumPy.
"""

from __future__ import annotations

import numpy as np


class Tiling:
    """Dyadic tiling; follows ``DyadicDecomposition`` (decomposition.py:218-541).

    Attributes mirror the reference: ``box_lengths`` (370-386), retained
    ``num_boxes`` per level (425-452), ``centres`` (454-541, meshgrid ``ij``
    order, inner cut-out by the previous level's bound).
    """

    def __init__(self, num_levels, N, num_boxes_levels, box_aspect_ratio,
                 boxes_per_dim_levels=None):
        self.num_levels = int(num_levels)
        self.N = tuple(int(n) for n in N)
        self.num_boxes_levels = tuple(int(x) for x in num_boxes_levels)
        self.aspect = tuple(int(x) for x in box_aspect_ratio)
        self.boxes_per_dim_levels = boxes_per_dim_levels
        self.ndim = len(self.N)
        n_ref = min(self.N)
        denom = self.num_boxes_levels[-1]
        # decomposition.py:370-386
        self.box_lengths = [
            n_ref // (denom * 2 ** (self.num_levels - 1 - lvl))
            for lvl in range(self.num_levels)
        ]
        self.num_boxes = self._retained_counts()
        self.centres = self._centres()
        self.boxes_cumsum = np.concatenate([[0], np.cumsum(self.num_boxes)]).astype(np.int64)
        self.total_num_boxes = int(self.boxes_cumsum[-1])

    def _outer_per_axis(self, lvl):
        # decomposition.py:390-423
        if self.boxes_per_dim_levels is not None:
            return np.asarray(self.boxes_per_dim_levels[lvl], dtype=np.int64)
        n = np.asarray(self.N, dtype=float)
        per = self.num_boxes_levels[lvl] * (n / n.min()) / np.asarray(self.aspect, float)
        return np.rint(per).astype(np.int64)

    def _retained_counts(self):
        # decomposition.py:425-452: retained = total - previous total / 2^d
        totals = [int(np.prod(self._outer_per_axis(l))) for l in range(self.num_levels)]
        out = [totals[0]]
        for lvl in range(1, self.num_levels):
            out.append(totals[lvl] - totals[lvl - 1] // (2 ** self.ndim))
        return np.asarray(out, dtype=np.int64)

    @staticmethod
    def _centres_1d(length, num_outer, scale):
        # decomposition.py:454-476
        step = int(length * scale)
        half = num_outer // 2
        return np.arange(-half * step + step // 2, half * step, step, dtype=np.int64)

    def _centres(self):
        # decomposition.py:478-541
        out = []
        bounds_prev = np.zeros(self.ndim, dtype=np.int64)
        for lvl in range(self.num_levels):
            step_len = self.box_lengths[lvl]
            per_axis = self._outer_per_axis(lvl)
            axes = [self._centres_1d(step_len, int(per_axis[a]), self.aspect[a])
                    for a in range(self.ndim)]
            mesh = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, self.ndim)
            if mesh.size > 0 and np.any(bounds_prev > 0):
                keep = ~np.all(np.abs(mesh) < bounds_prev, axis=1)
                mesh = mesh[keep]
            assert mesh.shape[0] == self.num_boxes[lvl]
            out.append(mesh)
            if mesh.shape[0] > 0:
                bounds_prev = np.max(np.abs(mesh), axis=0) + (step_len // 2)
        return np.concatenate(out, axis=0)

    def find_level(self, box):
        """``utils.find_level`` (coeff_index.py:90-108): searchsorted right."""
        return np.searchsorted(np.cumsum(self.num_boxes), box, side="right")

    @property
    def fourier_meshgrid(self):
        """decomposition.py:343-354: stacked ``arange(-n//2, n//2)`` (ij)."""
        axes = [np.arange(-n // 2, n // 2) for n in self.N]
        return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)

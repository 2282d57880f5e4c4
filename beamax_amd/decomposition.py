"""Dyadic Fourier tiling (host side, integer work).

API of ``beamax.decomposition.DyadicDecomposition`` (reference
``beamax/decomposition.py:218-541``): the constructor arguments, the public
attributes (``num_boxes_ndim``, ``centres_ndim``, ``total_num_boxes``,
``box_lengths``, ``fourier_meshgrid``, ``scaling``) and the ``ValueError``
messages of ``validate_params`` (``decomposition.py:89-215``) are kept, since
callers and tests match on them.

The tiling itself is built here as one integer table.  Level l has, along
axis a, ``outer[l, a]`` boxes of Fourier length ``step[l, a] = length[l] *
aspect[a]``, centred at ``step * (j - outer / 2) + step // 2``.  A level
keeps the boxes that reach outside the previous level's retained region (the
box at least as far out as that level's largest centre plus half its length
on some axis), so that the levels nest; the table is the concatenation of
the levels' lattices in C order of (j_0, ..., j_{d-1}), j_a < 2 (outer // 2).
The device gets the table once per plan.
"""

from __future__ import annotations

import math
from typing import Optional, Sequence, Tuple

import numpy as np


def _integral(value) -> bool:
    """A whole number given as an int or a finite float (not a bool)."""
    if isinstance(value, (bool, np.bool_)):
        return False
    if isinstance(value, (int, np.integer)):
        return True
    return isinstance(value, (float, np.floating)) and bool(np.isfinite(value)) \
        and float(value).is_integer()


def _ints(values: Sequence, message: str) -> Tuple[int, ...]:
    if any(not _integral(v) for v in values):
        raise ValueError(message)
    return tuple(int(v) for v in values)


def _fail_if(condition: bool, message: str) -> None:
    if condition:
        raise ValueError(message)


def _check_explicit_counts(bpl, num_levels: int, ndim: int) -> None:
    """Explicit boxes per axis and level (decomposition.py:39-86 messages)."""
    _fail_if(len(bpl) != num_levels,
             f"len(boxes_per_dim_levels)={len(bpl)} != num_levels={num_levels}.")
    for lvl, row in enumerate(bpl):
        _fail_if(len(row) != ndim,
                 f"Level {lvl}: len(boxes_per_dim_levels[lvl])={len(row)} != ndim={ndim}.")
        vals = _ints(row, f"Level {lvl}: boxes per axis must be integers; got {row}.")
        _fail_if(min(vals) <= 0, f"Level {lvl}: boxes per axis must be positive; got {row}.")
        _fail_if(any(v & 1 for v in vals),
                 f"Level {lvl}: boxes per axis must be even for symmetry; got {row}.")
    table = np.asarray([[int(v) for v in row] for row in bpl], dtype=np.int64)
    drops = np.flatnonzero(np.any(np.diff(table, axis=0) < 0, axis=0)) if num_levels > 1 else []
    if len(drops):
        ax = int(drops[0])
        raise ValueError("boxes_per_dim_levels must be nondecreasing per axis; axis "
                         f"{ax} has {table[:, ax].tolist()}.")


def validate_params(num_levels, N, num_boxes_levels, box_aspect_ratio,
                    boxes_per_dim_levels=None) -> None:
    """Raise ``ValueError`` (with ``validate_params``'s messages) unless the
    arguments define a dyadic tiling; the first failing condition in the
    order grid, aspect ratio, level counts, per-axis counts, level depth, box
    lengths is reported."""
    _fail_if(not isinstance(num_levels, int) or isinstance(num_levels, bool) or num_levels < 1,
             f"num_levels must be positive int; got {num_levels}.")
    _fail_if(not N, f"N entries must be integers; got {N}.")
    grid = _ints(N, f"N entries must be integers; got {N}.")
    _fail_if(min(grid) <= 0, f"N must be positive; got {grid}.")
    _fail_if(any(n & 1 for n in grid), f"N must be even along all axes; got {grid}.")
    ndim = len(grid)

    aspect = _ints(box_aspect_ratio,
                   f"box_aspect_ratio entries must be integers; got {box_aspect_ratio}.")
    for bad, msg in ((len(aspect) == 0, "box_aspect_ratio must be non-empty."),
                     (len(aspect) != ndim, f"len(box_aspect_ratio)={len(aspect)} != ndim={ndim}."),
                     (len(aspect) and min(aspect) <= 0,
                      f"box_aspect_ratio must be positive; got {aspect}."),
                     (1 not in aspect, "At least one aspect ratio component must equal 1."),
                     (ndim == 1 and aspect != (1,), "For 1D, box_aspect_ratio must be (1,).")):
        _fail_if(bool(bad), msg)

    counts = _ints(num_boxes_levels,
                   f"num_boxes_levels entries must be integers; got {num_boxes_levels}.")
    _fail_if(len(counts) != num_levels,
             f"len(num_boxes_levels)={len(counts)} != num_levels={num_levels}.")
    _fail_if(min(counts) <= 0, f"num_boxes_levels must be positive; got {counts}.")
    _fail_if(list(counts) != sorted(counts),
             f"num_boxes_levels must be non-decreasing with level; got {counts}.")

    n_ref = min(grid)
    if boxes_per_dim_levels is not None:
        _check_explicit_counts(boxes_per_dim_levels, num_levels, ndim)
        small = grid.index(n_ref)
        along = tuple(int(row[small]) for row in boxes_per_dim_levels)
        _fail_if(along != counts,
                 "boxes_per_dim_levels disagrees with num_boxes_levels on the smallest axis. "
                 f"Expected smallest-axis sequence {counts}, got {along}.")
    else:
        ratio = np.asarray(grid, float) / float(n_ref) / np.asarray(aspect, float)
        for boxes in counts:
            per_axis = boxes * ratio
            _fail_if(not np.allclose(per_axis, np.rint(per_axis), rtol=0.0, atol=1e-12),
                     "Derived boxes per axis are non-integral; provide a compatible "
                     "box_aspect_ratio or explicit boxes_per_dim_levels. "
                     f"Got counts {per_axis.tolist()} for {boxes} boxes.")

    base = counts[0]
    _fail_if(base > n_ref, f"Invalid base boxes per side: {base} for N_ref={n_ref}.")
    deepest = int(math.floor(math.log2(n_ref // base)) + 1)
    _fail_if(num_levels > deepest,
             f"num_levels must be in [1, {deepest}] (N_ref={n_ref}, base={base}).")
    lengths = [int(v) for v in _level_lengths(n_ref, counts[-1], num_levels)]
    _fail_if(min(lengths) < 2,
             "Every box length must be at least 2 Fourier samples so the "
             f"Gaussian width is positive; computed {lengths} for N_ref={n_ref}.")
    _fail_if(any(n_ref % v for v in lengths),
             f"Box lengths must divide N_ref={n_ref}; computed {lengths}.")


def _level_lengths(n_ref: int, finest_count: int, num_levels: int) -> np.ndarray:
    """Box length along the smallest axis per level: the finest level has
    ``finest_count`` boxes, each coarser level boxes twice as long."""
    halvings = np.arange(num_levels - 1, -1, -1, dtype=np.int64)
    return n_ref // (finest_count << halvings)


def _outer_counts(N, counts, aspect, explicit) -> np.ndarray:
    """(levels, d) boxes per axis of each level's full lattice."""
    if explicit is not None:
        return np.asarray([[int(v) for v in row] for row in explicit], dtype=np.int64)
    n = np.asarray(N, dtype=np.int64)
    # counts * (N / min N) / aspect; validated integral
    num = np.asarray(counts, dtype=np.int64)[:, None] * n[None, :]
    den = int(n.min()) * np.asarray(aspect, dtype=np.int64)[None, :]
    return num // den


def _centre_table(lengths, outer, aspect):
    """Centres of every retained box, (total, d) int32, and the retained
    count per level.  One vectorised pass per level over its lattice."""
    d = outer.shape[1]
    asp = np.asarray(aspect, dtype=np.int64)
    table, kept = [], []
    reach = None                      # previous level's retained half-extent per axis
    for lvl, length in enumerate(lengths):
        step = int(length) * asp
        half = outer[lvl] // 2        # an odd count leaves its last box out
        j = np.indices(tuple(int(v) for v in 2 * half), dtype=np.int64).reshape(d, -1).T
        cen = step * (j - half) + step // 2
        if reach is not None:
            cen = cen[np.any(np.abs(cen) >= reach, axis=1)]
        table.append(cen)
        kept.append(cen.shape[0])
        if cen.shape[0]:
            reach = np.abs(cen).max(axis=0) + int(length) // 2
    return np.concatenate(table, axis=0).astype(np.int32), np.asarray(kept, dtype=np.int64)


class DyadicDecomposition:
    """Multi-level dyadic tiling of Fourier space (decomposition.py:218-541).

    Attributes: ``num_boxes_ndim`` (retained boxes per level), ``centres_ndim``
    (total_boxes, d) int32 Fourier-index centres, ``total_num_boxes``,
    ``box_lengths``, ``fourier_meshgrid``, ``scaling``.
    """

    def __init__(self, num_levels: int, N: Tuple[int, ...], num_boxes_levels: Tuple[int, ...],
                 box_aspect_ratio: Tuple[int, ...],
                 boxes_per_dim_levels: Optional[Tuple[Tuple[int, ...], ...]] = None):
        validate_params(num_levels, N, num_boxes_levels, box_aspect_ratio, boxes_per_dim_levels)
        self.num_levels = int(num_levels)
        self.N = tuple(int(x) for x in N)
        self.num_boxes_levels = tuple(int(x) for x in num_boxes_levels)
        self.box_aspect_ratio = tuple(int(x) for x in box_aspect_ratio)
        self.boxes_per_dim_levels = boxes_per_dim_levels
        outer = _outer_counts(self.N, self.num_boxes_levels, self.box_aspect_ratio,
                              boxes_per_dim_levels)
        centres, kept = _centre_table(self.box_lengths, outer, self.box_aspect_ratio)
        # a level nests inside the next one only if its lattice, halved per
        # axis, is exactly the part the next level drops
        expected = np.prod(outer, axis=1)
        expected[1:] -= np.prod(outer[:-1], axis=1) // (2 ** self.ndim)
        for lvl in np.flatnonzero(kept != expected):
            raise ValueError(f"Level {lvl}: expected {int(expected[lvl])} centres, got "
                             f"{int(kept[lvl])}. (per_axis_counts={outer[lvl].tolist()})")
        self.num_boxes_ndim = expected.astype(np.int32)
        self.centres_ndim = centres
        self.total_num_boxes = int(self.num_boxes_ndim.sum())

    @property
    def ndim(self) -> int:
        return len(self.N)

    @property
    def num_boxes_ndim_cumsum(self):
        return np.cumsum(self.num_boxes_ndim)

    @property
    def fourier_meshgrid(self):
        axes = [np.arange(-n // 2, n // 2, dtype=np.int32) for n in self.N]
        return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)

    @property
    def scaling(self):
        n = np.asarray(self.N, float)
        return n / n.min() / np.asarray(self.box_aspect_ratio, float)

    @property
    def box_lengths(self):
        """Box length along the smallest axis per level (decomposition.py:370-386)."""
        return _level_lengths(min(self.N), self.num_boxes_levels[-1],
                              self.num_levels).astype(np.int32)

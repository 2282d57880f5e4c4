"""Oracle: MSWPT analysis/synthesis in FP64 NumPy (test infrastructure only).

Restates ``beamax/transforms.py`` and ``beamax/utils/fft.py`` of the reference.
"""

from __future__ import annotations

import numpy as np

from .decomposition import Tiling


def unitary_fft(x):
    """utils/fft.py:5-20: fftshift(fftn(x, ortho))."""
    return np.fft.fftshift(np.fft.fftn(x, norm="ortho"))


def unitary_ifft(x):
    """utils/fft.py:23-38: ifftn(ifftshift(x), ortho)."""
    return np.fft.ifftn(np.fft.ifftshift(x), norm="ortho")


def windowed_gaussian(centre, tiling: Tiling, level, redundancy, windowing):
    """transforms.py:36-96 on the full centred Fourier grid.

    sigma = (bl // 2) * aspect; diff = ((m - c + N/2) mod N) - N/2 (float);
    g = exp(-sum (diff/sigma)^2), times the rectangular window.
    """
    bl = tiling.box_lengths[level]
    N = np.asarray(tiling.N, dtype=float)
    asp = np.asarray(tiling.aspect)
    sigma = (bl // 2) * asp
    mg = tiling.fourier_meshgrid
    diff = np.mod(mg - np.asarray(centre) + N / 2, N) - N / 2
    g = np.exp(-np.sum(np.square(diff / sigma), axis=-1))
    if windowing == "none":
        return g
    ws = bl * asp * (redundancy / 2)
    if windowing == "rectangular":
        win = np.all((-ws <= diff) & (diff < ws), axis=-1)
    elif windowing == "rectangular_mirror":
        win = np.all(np.abs(diff) < ws, axis=-1)
    else:
        raise ValueError(windowing)
    return g * win.astype(float)


def filter_idx(box, tiling, redundancy, windowing):
    """transforms.py:99-138 (single_filter_idx)."""
    lvl = int(tiling.find_level(box))
    return windowed_gaussian(tiling.centres[box], tiling, lvl, redundancy, windowing)


def frame_phase(tiling: Tiling, box, k, redundancy):
    """transforms.py:193-226: unit phase from the parity roll of the patch.

    Uses ``centre + N//2`` for the parity index (as the reference does here).
    """
    box = np.asarray(box)
    k = np.asarray(k)
    lvl = tiling.find_level(box)
    bl = np.asarray(tiling.box_lengths)[lvl]
    if np.ndim(bl) > 0:
        bl = bl[..., None]
    support = bl * redundancy * np.asarray(tiling.aspect)
    half = support // 2
    centre = tiling.centres[box]
    parity = (centre + np.asarray(tiling.N) // 2) // half
    sign = 1 - 2 * (parity & 1)
    rolls = sign * (half // 2)
    cyc = np.sum((centre - rolls) * k / support, axis=-1)
    cyc = np.remainder(cyc, 1.0)
    return np.exp(2 * np.pi * 1j * cyc)


def compute_frames(tiling, box, k, redundancy, windowing="rectangular"):
    """transforms.py:309-354: Fourier-domain frame atom of (box, k)."""
    lvl = int(tiling.find_level(box))
    bl = tiling.box_lengths[lvl]
    support = bl * redundancy * np.asarray(tiling.aspect)
    g = filter_idx(box, tiling, redundancy, windowing)
    scaled = np.einsum("...d,d->...", tiling.fourier_meshgrid, np.asarray(k) / support)
    norm = np.prod(support) ** (-0.5)
    ph = frame_phase(tiling, box, np.asarray(k), redundancy)
    return norm * ph * np.exp(-2 * np.pi * 1j * scaled) * g


def sum_gsquare_windowed(tiling: Tiling, redundancy, windowing):
    """transforms.py:511-555: sum over every box, in box order, of the
    squared windowed filter.  Each filter is evaluated only where its window
    can be non-zero (offsets |o| <= ceil(ws) + 1 from its centre on each axis,
    wrapped): elsewhere it is exactly zero, so the sum equals the full-grid
    sum of ``windowed_gaussian`` squares bit for bit (the same operations on
    the same values, in the same order)."""
    N = np.asarray(tiling.N)
    Nf = N.astype(float)
    d = tiling.ndim
    asp = np.asarray(tiling.aspect)
    sg = np.zeros(tiling.N)
    for b in range(tiling.total_num_boxes):
        lvl = int(tiling.find_level(b))
        bl = tiling.box_lengths[lvl]
        sigma = (bl // 2) * asp
        ws = bl * asp * (redundancy / 2)
        centre = tiling.centres[b]
        idx, sq, win = [], [], []
        for a in range(d):
            r = int(np.ceil(ws[a])) + 1
            m = np.unique((centre[a] + N[a] // 2 + np.arange(-r, r + 1)) % N[a])
            mg = m - N[a] // 2                      # fourier_meshgrid value at index m
            diff = np.mod(mg - centre[a] + Nf[a] / 2, Nf[a]) - Nf[a] / 2
            shape = [1] * d
            shape[a] = m.size
            idx.append(m)
            sq.append(np.square(diff / sigma[a]).reshape(shape))
            if windowing == "rectangular":
                w = (-ws[a] <= diff) & (diff < ws[a])
            elif windowing == "rectangular_mirror":
                w = np.abs(diff) < ws[a]
            else:
                raise ValueError(windowing)
            win.append(w.reshape(shape))
        tot = sq[0]
        for a in range(1, d):
            tot = tot + sq[a]
        allw = win[0]
        for a in range(1, d):
            allw = allw & win[a]
        g = np.exp(-tot) * allw.astype(float)
        ix = np.ix_(*idx)
        sg[ix] = sg[ix] + np.square(g)
    return sg


def _extract_centered_box(arr, shape, centre):
    """utils/arrays.py:237-268: periodic extraction of a centred box."""
    N = arr.shape
    out = arr
    for ax in range(arr.ndim):
        start = (centre[ax] - shape[ax] // 2) % N[ax]
        idx = (start + np.arange(shape[ax])) % N[ax]
        out = np.take(out, idx, axis=ax)
    return out


class MSWPT:
    """transforms.py:357-850 restated; coefficient layout identical.

    Flat layout: levels concatenated; within a level (box, k_1..k_d) C order.
    """

    def __init__(self, tiling: Tiling, redundancy=2, windowing="rectangular_mirror"):
        if redundancy not in (1, 2):
            raise ValueError("redundancy")
        if windowing == "none":
            raise ValueError("MSWPT.forward does not support windowing='none'")
        self.tiling = tiling
        self.redundancy = redundancy
        self.windowing = windowing
        d = tiling.ndim
        # transforms.py:511-555: sum of squares over every box, box order.
        sg = sum_gsquare_windowed(tiling, redundancy, windowing)
        if not np.all(sg > 0):
            raise ValueError("uncovered Fourier bins")
        self.sum_gsquare = sg
        # utils/coeff_index.py:150-180 (coefficient shapes)
        self.coeff_shapes = []
        self.support_shapes = []
        self.box_shapes = []
        for lvl in range(tiling.num_levels):
            bl = tiling.box_lengths[lvl]
            cs = tuple(int(a * redundancy * bl) for a in tiling.aspect)
            self.coeff_shapes.append((int(tiling.num_boxes[lvl]),) + cs)
            bs = tuple(int(bl * a * redundancy / 2) for a in tiling.aspect)
            self.box_shapes.append(bs)
            self.support_shapes.append(tuple(2 * x for x in bs))
        self.coeffs_cumsum = np.concatenate(
            [[0], np.cumsum([int(np.prod(s)) for s in self.coeff_shapes])]).astype(np.int64)
        self.total_coeffs = int(self.coeffs_cumsum[-1])
        seg = np.diff(self.coeffs_cumsum)
        # transforms.py:502-508
        self.half_mask = np.concatenate(
            [np.concatenate([np.ones(L // 2), np.zeros(L - L // 2)]) for L in seg])
        # transforms.py:557-589: packed filter of the first box of each level
        N = np.asarray(tiling.N)
        self.gpacked = []
        for lvl in range(tiling.num_levels):
            start = int(tiling.boxes_cumsum[lvl])
            g = filter_idx(start, tiling, redundancy, windowing)
            c = tiling.centres[start] + N // 2
            self.gpacked.append(_extract_centered_box(g, self.support_shapes[lvl], c))
        assert d == len(tiling.N)

    def _rolls(self, centres_shifted, lvl):
        # transforms.py:665-672: parity from (centre + N//2 + N//2) // box_shape
        N = np.asarray(self.tiling.N)
        bs = np.asarray(self.box_shapes[lvl])
        ri = (centres_shifted + N // 2) // bs
        sign = 1 - 2 * (ri & 1)
        return sign * (bs // 2)

    def forward(self, data, input_type="spatial"):
        """transforms.py:696-733 and 591-694."""
        t = self.tiling
        d = t.ndim
        F = unitary_fft(data) if input_type == "spatial" else np.asarray(data, complex)
        F = F / self.sum_gsquare
        N = np.asarray(t.N)
        out = np.zeros(self.total_coeffs, dtype=complex)
        for lvl in range(t.num_levels):
            b0, b1 = int(t.boxes_cumsum[lvl]), int(t.boxes_cumsum[lvl + 1])
            cen = t.centres[b0:b1] + N // 2
            S = self.support_shapes[lvl]
            # periodic gather of every box's patch at once
            idx = [((cen[:, a, None] - S[a] // 2) % N[a] + np.arange(S[a])) % N[a]
                   for a in range(d)]
            if d == 1:
                patch = F[idx[0]]
            elif d == 2:
                patch = F[idx[0][:, :, None], idx[1][:, None, :]]
            else:
                patch = F[idx[0][:, :, None, None], idx[1][:, None, :, None],
                          idx[2][:, None, None, :]]
            sf = patch * self.gpacked[lvl][None]
            rolls = self._rolls(cen, lvl)
            # np.roll per box: out[j] = in[(j - r) mod S]
            for a in range(d):
                j = (np.arange(S[a])[None, :] - rolls[:, a, None]) % S[a]
                shape = [len(cen)] + [1] * d
                shape[1 + a] = S[a]
                sf = np.take_along_axis(sf, j.reshape(shape), axis=1 + a)
            axes = tuple(range(1, d + 1))
            coeff = np.fft.ifftn(np.fft.ifftshift(sf, axes=axes), axes=axes, norm="ortho")
            out[self.coeffs_cumsum[lvl]:self.coeffs_cumsum[lvl + 1]] = coeff.ravel()
        return out

    def inverse(self, coeffs, output_type="spatial"):
        """transforms.py:735-850: g * roll(FFT(c), -r) scatter-added at the box."""
        t = self.tiling
        d = t.ndim
        N = np.asarray(t.N)
        ft = np.zeros(t.N, dtype=complex)
        for lvl in range(t.num_levels):
            b0, b1 = int(t.boxes_cumsum[lvl]), int(t.boxes_cumsum[lvl + 1])
            S = self.support_shapes[lvl]
            c_lvl = coeffs[self.coeffs_cumsum[lvl]:self.coeffs_cumsum[lvl + 1]]
            c_lvl = c_lvl.reshape((b1 - b0,) + tuple(S))
            bl = np.asarray(self.box_shapes[lvl])
            for i in range(b1 - b0):
                centre = t.centres[b0 + i]
                ri = (centre + N // 2) // bl
                rolls = (1 - 2 * (ri & 1)) * (bl // 2)
                fh = unitary_fft(c_lvl[i])
                for a in range(d):
                    fh = np.take(fh, (np.arange(S[a]) + rolls[a]) % S[a], axis=a)
                contrib = self.gpacked[lvl] * fh
                c0 = (centre + N // 2) % N
                starts = (c0 - np.asarray(S) // 2) % N
                grids = np.meshgrid(*[(starts[a] + np.arange(S[a])) % N[a] for a in range(d)],
                                    indexing="ij")
                np.add.at(ft, tuple(grids), contrib)
        return unitary_ifft(ft) if output_type == "spatial" else ft

    def adjoint(self, coeffs):
        """W^H: the adjoint of ``forward`` for the inner product Re<., .>.

        ``forward`` is unitary_fft, division by sum_gsquare, a gather of each
        box's patch, the window, a roll and a unitary inverse box FFT; W^H is
        the reverse chain of adjoints, i.e. the synthesis of ``inverse`` in the
        Fourier domain divided by sum_gsquare, then unitary_ifft.
        """
        return unitary_ifft(self.inverse(coeffs, output_type="fourier") / self.sum_gsquare)

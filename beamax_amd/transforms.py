"""MSWPT: multiscale wave-packet transform on the GPU.

API of ``beamax.transforms.MSWPT`` (reference ``beamax/transforms.py:357-912``).
The plan tables (sum of squared filters, packed first-box filters) are built on
the device by ``bh_mswpt_setup``; ``forward`` / ``inverse`` run the HIP kernels
in ``csrc/bh_mswpt.hip`` and ``csrc/bh_fft.hip`` in FP64 on torch device
tensors.  ``compute_frames`` / ``compute_frame_phase`` are small host helpers
(NumPy) used to build test inputs, as in the reference README example.
"""

from __future__ import annotations

import ctypes as C
from typing import Tuple

import numpy as np

from . import _lib
from .decomposition import DyadicDecomposition


def _validate_transform_configuration(redundancy: int, windowing: str) -> None:
    """transforms.py:22-33."""
    if redundancy not in (1, 2):
        raise ValueError(f"redundancy must be 1 or 2; got {redundancy}.")
    if windowing not in ("none", "rectangular", "rectangular_mirror"):
        raise ValueError(f"Unknown windowing mode {windowing!r}.")
    if redundancy == 1 and windowing == "rectangular_mirror":
        raise ValueError(
            "redundancy=1 with windowing='rectangular_mirror' leaves "
            "uncovered Fourier bins, so the canonical-dual analysis is "
            "undefined; use windowing='rectangular' or redundancy=2.")


def _default_device():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("beamax_amd needs a ROCm GPU (torch.cuda is not available)")
    return torch.device("cuda", torch.cuda.current_device())


def _level_of(decomp: DyadicDecomposition, box):
    return np.searchsorted(decomp.num_boxes_ndim_cumsum, box, side="right")


def compute_windowed_gaussian(centre, meshgrid, box_length, box_aspect_ratio, domain_length,
                              redundancy, windowing):
    """Host evaluation of one tile filter (transforms.py:36-96)."""
    asp = np.asarray(box_aspect_ratio)
    L = np.asarray(domain_length, dtype=float)
    sigma = (int(box_length) // 2) * asp
    diff = np.mod(np.asarray(meshgrid) - np.asarray(centre) + L / 2, L) - L / 2
    g = np.exp(-np.sum(np.square(diff / sigma), axis=-1))
    if windowing == "none":
        return g
    ws = int(box_length) * asp * (redundancy / 2)
    if windowing == "rectangular":
        win = np.all((-ws <= diff) & (diff < ws), axis=-1)
    elif windowing == "rectangular_mirror":
        win = np.all(np.abs(diff) < ws, axis=-1)
    else:
        raise ValueError(f"Invalid windowing function: {windowing}")
    return g * win.astype(float)


def compute_frame_phase(decomp: DyadicDecomposition, boxidx, k, redundancy):
    """Unit phase of the parity-rolled patch (transforms.py:193-226)."""
    boxidx = np.asarray(boxidx)
    lvl = _level_of(decomp, boxidx)
    bl = decomp.box_lengths[lvl]
    if np.ndim(bl) > 0:
        bl = bl[..., None]
    support = bl * redundancy * np.asarray(decomp.box_aspect_ratio)
    half = support // 2
    centre = decomp.centres_ndim[boxidx].astype(np.int64)
    parity = (centre + np.asarray(decomp.N) // 2) // half
    rolls = (1 - 2 * (parity & 1)) * (half // 2)
    cyc = np.remainder(np.sum((centre - rolls) * np.asarray(k) / support, axis=-1), 1.0)
    return np.exp(2 * np.pi * 1j * cyc)


def compute_frames(dyadic_decomp: DyadicDecomposition, boxidx: int, k, fourier_space,
                   redundancy: int, windowing: str = "rectangular"):
    """Fourier-domain frame atom of box `boxidx`, modulation `k` (transforms.py:309-354)."""
    lvl = int(_level_of(dyadic_decomp, boxidx))
    bl = int(dyadic_decomp.box_lengths[lvl])
    support = bl * redundancy * np.asarray(dyadic_decomp.box_aspect_ratio)
    g = compute_windowed_gaussian(dyadic_decomp.centres_ndim[boxidx], dyadic_decomp.fourier_meshgrid,
                                  bl, dyadic_decomp.box_aspect_ratio, dyadic_decomp.N, redundancy,
                                  windowing)
    scaled = np.einsum("...d,d->...", np.asarray(fourier_space), np.asarray(k) / support)
    norm = np.prod(support) ** (-0.5)
    ph = compute_frame_phase(dyadic_decomp, boxidx, np.asarray(k), redundancy)
    return norm * ph * np.exp(-2 * np.pi * 1j * scaled) * g


def unitary_ifft_host(x):
    """ifftn(ifftshift(x), ortho) on the host (utils/fft.py:23-38)."""
    return np.fft.ifftn(np.fft.ifftshift(x), norm="ortho")


class MSWPT:
    """Multiscale wave-packet transform (transforms.py:357-850), FP64 on the GPU.

    Attributes follow the reference: ``total_coeffs``, ``coeffs_cumsum``,
    ``coeff_shapes``, ``boxes_cumsum``, ``_support_shapes``, ``_box_shapes``,
    ``_half_mask`` (NumPy), ``sum_gsquare`` and ``gfilts_packed`` (device).
    """

    def __init__(self, dyadic_decomp: DyadicDecomposition, redundancy: int = 2,
                 windowing: str = "rectangular", device=None):
        import torch
        _validate_transform_configuration(redundancy, windowing)
        self.dyadic_decomp = dyadic_decomp
        self.redundancy = int(redundancy)
        self.windowing = windowing
        d = dyadic_decomp.ndim
        if dyadic_decomp.num_levels > _lib.BH_MAX_LEVELS:
            raise ValueError(f"at most {_lib.BH_MAX_LEVELS} levels are supported")
        self.boxes_cumsum = tuple(int(x) for x in
                                  np.concatenate([[0], dyadic_decomp.num_boxes_ndim_cumsum]))
        bls = [int(x) for x in dyadic_decomp.box_lengths]
        self.coeff_shapes = tuple(
            (int(dyadic_decomp.num_boxes_ndim[l]),) +
            tuple(int(a * redundancy * bls[l]) for a in dyadic_decomp.box_aspect_ratio)
            for l in range(dyadic_decomp.num_levels))
        cc = [0]
        for s in self.coeff_shapes:
            cc.append(cc[-1] + int(np.prod(s)))
        self.coeffs_cumsum = tuple(cc)
        self.total_coeffs = int(cc[-1])
        self.max_box_length = int(bls[-1] * max(dyadic_decomp.box_aspect_ratio) * (redundancy / 2))
        self._box_shapes = [tuple(int(bls[l] * a * redundancy / 2) for a in dyadic_decomp.box_aspect_ratio)
                            for l in range(dyadic_decomp.num_levels)]
        self._support_shapes = [tuple(2 * x for x in bs) for bs in self._box_shapes]
        seg = np.diff(np.asarray(cc))
        self._half_mask = np.concatenate(
            [np.concatenate([np.ones(L // 2), np.zeros(L - L // 2)]) for L in seg])
        self.device = torch.device(device) if device is not None else _default_device()
        if windowing == "none":
            # synthesis-only configuration (transforms.py:367-371); no fast analysis
            self.sum_gsquare = None
            self.gfilts_packed = None
            self._tiling = None
            return
        self._centres_dev = torch.as_tensor(np.ascontiguousarray(dyadic_decomp.centres_ndim,
                                                                 dtype=np.int32),
                                            device=self.device)
        self._tiling = self._make_tiling()
        ntot = int(np.prod(dyadic_decomp.N))
        npack = sum(int(np.prod(s)) for s in self._support_shapes)
        self.sum_gsquare = torch.empty(dyadic_decomp.N, dtype=torch.float64, device=self.device)
        self.gfilts_packed = torch.empty(npack, dtype=torch.float64, device=self.device)
        _lib.call("bh_mswpt_setup", C.byref(self._tiling), _lib.ptr(self.sum_gsquare),
                  _lib.ptr(self.gfilts_packed), _lib.stream_handle())
        ok = bool(torch.all(torch.isfinite(self.sum_gsquare) & (self.sum_gsquare > 0)).item())
        if not ok:
            raise ValueError("The dyadic decomposition leaves uncovered Fourier bins; "
                             "choose compatible grid, box counts, and aspect ratios.")
        self._ws_bytes = int(_lib.load().bh_mswpt_workspace_size(C.byref(self._tiling)))
        assert ntot == self.sum_gsquare.numel()

    def _make_tiling(self) -> "_lib.Tiling":
        dd = self.dyadic_decomp
        t = _lib.Tiling()
        t.ndim = dd.ndim
        t.num_levels = dd.num_levels
        t.redundancy = self.redundancy
        t.windowing = _lib.WINDOW[self.windowing]
        for a in range(dd.ndim):
            t.N[a] = dd.N[a]
            t.aspect[a] = dd.box_aspect_ratio[a]
        for l in range(dd.num_levels):
            t.box_length[l] = int(dd.box_lengths[l])
            for a in range(dd.ndim):
                t.support[l][a] = self.coeff_shapes[l][1 + a]
        for l in range(dd.num_levels + 1):
            t.boxes_cumsum[l] = self.boxes_cumsum[l]
            t.coeffs_cumsum[l] = self.coeffs_cumsum[l]
        for l in range(dd.num_levels + 1, _lib.BH_MAX_LEVELS + 1):
            t.boxes_cumsum[l] = self.boxes_cumsum[-1]
            t.coeffs_cumsum[l] = self.coeffs_cumsum[-1]
        t.centres = self._centres_dev.data_ptr()
        return t

    @property
    def tiling(self):
        return self._tiling

    @property
    def half_frame_capacity(self) -> int:
        seg = np.diff(np.asarray(self.coeffs_cumsum))
        return int(np.sum(seg // 2))

    def _as_device(self, data, dtype):
        import torch
        if isinstance(data, torch.Tensor):
            t = data.to(device=self.device, dtype=dtype)
        else:
            t = torch.as_tensor(np.asarray(data), device=self.device).to(dtype)
        return t.contiguous()

    def forward(self, data, input_type: str = "spatial", *, scale: float = 1.0,
                half_mask: bool = False, out=None, ws=None):
        """Flat complex128 coefficients (total_coeffs,) of `data` (transforms.py:696-733).

        ``scale`` and ``half_mask`` fold the ``compute_coefficients`` 'pos_only'
        epilogue (0.5 * a * half_mask) into the same launch sequence.
        """
        import torch
        if self.windowing == "none":
            raise ValueError(
                "MSWPT.forward does not support windowing='none': the fast "
                "local-patch implementation truncates an otherwise global "
                "Gaussian, so it cannot provide an exact analysis/synthesis "
                "pair. Use 'rectangular' or 'rectangular_mirror'.")
        if input_type not in ("spatial", "fourier"):
            raise ValueError("Invalid conversion.")
        shape = tuple(data.shape)
        if shape != self.dyadic_decomp.N:
            raise ValueError(f"data must have shape {self.dyadic_decomp.N}; got {shape}.")
        is_complex = (data.is_complex() if isinstance(data, torch.Tensor)
                      else np.iscomplexobj(np.asarray(data)))
        if input_type == "fourier":
            x = self._as_device(data, torch.complex128)
            kind = _lib.INPUT_COMPLEX_FOURIER
        elif is_complex:
            x = self._as_device(data, torch.complex128)
            kind = _lib.INPUT_COMPLEX_SPATIAL
        else:
            x = self._as_device(data, torch.float64)
            kind = _lib.INPUT_REAL_SPATIAL
        if out is None:
            out = torch.empty(self.total_coeffs, dtype=torch.complex128, device=self.device)
        if ws is None:
            ws = _lib.workspace(self._ws_bytes, self.device)
        _lib.call("bh_mswpt_forward", C.byref(self._tiling), _lib.ptr(self.sum_gsquare),
                  _lib.ptr(self.gfilts_packed), _lib.ptr(x), kind, float(scale), int(half_mask),
                  _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream_handle())
        return out

    def inverse(self, coeffs, output_type: str = "spatial"):
        """Fast synthesis, the exact inverse of ``forward`` (transforms.py:735-850)."""
        import torch
        if self._tiling is None:
            raise NotImplementedError("synthesis with windowing='none' is not built on the GPU")
        if output_type not in ("spatial", "fourier"):
            raise ValueError("Invalid conversion.")
        c = self._as_device(coeffs, torch.complex128)
        if c.numel() != self.total_coeffs:
            raise ValueError(f"coeffs must have {self.total_coeffs} entries; got {c.numel()}.")
        out = torch.empty(self.dyadic_decomp.N, dtype=torch.complex128, device=self.device)
        ws = _lib.workspace(self._ws_bytes, self.device)
        _lib.call("bh_mswpt_inverse", C.byref(self._tiling), _lib.ptr(self.gfilts_packed),
                  _lib.ptr(c), int(output_type == "fourier"), _lib.ptr(out), _lib.ptr(ws),
                  ws.numel(), _lib.stream_handle())
        return out

    def adjoint(self, coeffs, *, real: bool = True, out=None, ws=None):
        """W^H coeffs, the adjoint of ``forward`` (spatial input, no scale or mask).

        ``forward`` divides the spectrum by the sum of squared filters before
        windowing, so W^H is the synthesis of ``inverse`` divided by that sum
        before the inverse FFT (W^H W != I; W^-1 = ``inverse``).  ``real``
        returns the real part, the transpose for real spatial input.
        """
        import torch
        if self._tiling is None:
            raise NotImplementedError("the adjoint with windowing='none' is not built on the GPU")
        c = self._as_device(coeffs, torch.complex128)
        if c.numel() != self.total_coeffs:
            raise ValueError(f"coeffs must have {self.total_coeffs} entries; got {c.numel()}.")
        if out is None:
            out = torch.empty(self.dyadic_decomp.N,
                              dtype=torch.float64 if real else torch.complex128,
                              device=self.device)
        if ws is None:
            ws = _lib.workspace(self._ws_bytes, self.device)
        _lib.call("bh_mswpt_adjoint", C.byref(self._tiling), _lib.ptr(self.sum_gsquare),
                  _lib.ptr(self.gfilts_packed), _lib.ptr(c), int(real), _lib.ptr(out),
                  _lib.ptr(ws), ws.numel(), _lib.stream_handle())
        return out

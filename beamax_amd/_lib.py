"""ctypes binding of ``libbeamax_amd.so`` (the C-ABI in ``include/beamax_amd.h``).

The library is built in-tree by ``beamax_amd/build.py``.  Loading fails loudly
when it is missing: there is no CPU fallback for any operator.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libbeamax_amd.so"

BH_MAX_DIM = 3
BH_MAX_LEVELS = 12
WINDOW = {"rectangular": 1, "rectangular_mirror": 2}
INPUT_REAL_SPATIAL, INPUT_COMPLEX_SPATIAL, INPUT_COMPLEX_FOURIER = 0, 1, 2
PRECISION = {"fp64": 0, "fp32_split": 1}


class Tiling(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("num_levels", C.c_int32),
        ("redundancy", C.c_int32),
        ("windowing", C.c_int32),
        ("N", C.c_int64 * BH_MAX_DIM),
        ("aspect", C.c_int32 * BH_MAX_DIM),
        ("box_length", C.c_int32 * BH_MAX_LEVELS),
        ("support", (C.c_int32 * BH_MAX_DIM) * BH_MAX_LEVELS),
        ("boxes_cumsum", C.c_int64 * (BH_MAX_LEVELS + 1)),
        ("coeffs_cumsum", C.c_int64 * (BH_MAX_LEVELS + 1)),
        ("centres", C.c_void_p),
    ]


class FwdBeams(C.Structure):
    _fields_ = [("count", C.c_int64), ("x0", C.c_void_p), ("p0", C.c_void_p),
                ("omega", C.c_void_p), ("alpha", C.c_void_p), ("a0", C.c_void_p)]


class StaticBeams(C.Structure):
    _fields_ = [("count", C.c_int64), ("x", C.c_void_p), ("p", C.c_void_p),
                ("M", C.c_void_p), ("A", C.c_void_p), ("omega", C.c_void_p)]


class Segs(C.Structure):
    _fields_ = [("nseg", C.c_int64), ("ngroups", C.c_int64), ("S", C.c_int64),
                ("seg_p0", C.c_void_p), ("seg_v", C.c_void_p), ("seg_n", C.c_void_p),
                ("seg_g0", C.c_void_p), ("grp_seg", C.c_void_p), ("pts", C.c_void_p),
                ("col", C.c_void_p)]


class Speed(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("method", C.c_int32), ("boundary", C.c_int32),
                ("N", C.c_int64 * 3), ("spacing", C.c_double * 3), ("origin", C.c_double * 3),
                ("values", C.c_void_p), ("lam", C.c_double)]


class BatchSolve(C.Structure):
    _fields_ = [("beams", FwdBeams), ("c", C.c_double), ("ts", C.c_void_p), ("nt", C.c_int64),
                ("seg", Segs), ("out", C.c_void_p), ("win", C.c_void_p)]


_P = C.c_void_p
_I64 = C.c_int64
_SZ = C.c_size_t

# name -> (restype, argtypes); mirrors include/beamax_amd.h
SIGNATURES = {
    "bh_last_error": (C.c_char_p, []),
    "bh_version": (C.c_int, []),
    "bh_device_status": (C.c_int, [C.c_int, C.POINTER(C.c_uint32), _P]),
    "bh_fft_workspace_size": (_SZ, [_I64, _I64]),
    "bh_fft": (C.c_int, [_P, _I64, _I64, _I64, _I64, _I64, _I64, C.c_int, C.c_double, _P, _SZ, _P]),
    "bh_mswpt_setup": (C.c_int, [C.POINTER(Tiling), _P, _P, _P]),
    "bh_mswpt_workspace_size": (_SZ, [C.POINTER(Tiling)]),
    "bh_mswpt_forward": (C.c_int, [C.POINTER(Tiling), _P, _P, _P, C.c_int, C.c_double, C.c_int,
                                   _P, _P, _SZ, _P]),
    "bh_mswpt_inverse": (C.c_int, [C.POINTER(Tiling), _P, _P, C.c_int, _P, _P, _SZ, _P]),
    "bh_topn_workspace_size": (_SZ, [_I64, _I64]),
    "bh_threshold_workspace_size": (_SZ, [_I64]),
    "bh_threshold": (C.c_int, [_P, _I64, C.c_int, C.c_double, C.POINTER(Tiling), _P, _P, _P, _P, _SZ,
                               _P]),
    "bh_topn": (C.c_int, [_P, _I64, _I64, _P, _P, _P, _SZ, _P]),
    "bh_topn_noalloc": (C.c_int, [_P, _I64, _I64, _P, _P, _P]),
    "bh_forward_params": (C.c_int, [C.POINTER(Tiling), C.POINTER(C.c_double), _P, _P, _I64,
                                    FwdBeams, _P]),
    "bh_forward_sum": (C.c_int, [C.POINTER(FwdBeams), C.c_int, C.c_int, C.c_double, _P, _I64, _P,
                                 _I64, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int, _P,
                                 _P]),
    "bh_solve_hom_diag": (C.c_int, [_I64, C.c_int, _P, _P, _P, _P, _P, C.c_double, _P, _I64, _P,
                                    _P, _P, _P, _P]),
    "bh_solve_hom_TR": (C.c_int, [_I64, C.c_int, _P, _P, _P, _P, _P, C.c_double, _P, _I64, _P,
                                  _P, _P, _P, _P]),
    "bh_adjoint_source_workspace_size": (_SZ, [_I64, _I64]),
    "bh_adjoint_source": (C.c_int, [_P, _I64, _I64, C.c_double, _P, C.c_int, _P, _P, _P, _SZ, _P]),
    "bh_fourier_crop_time_workspace_size": (_SZ, [_I64, _I64, _I64]),
    "bh_fourier_crop_time": (C.c_int, [_P, _I64, _I64, _P, _P, _I64, _I64, _P, _P, _SZ, _P]),
    "bh_adjoint_params": (C.c_int, [C.POINTER(Tiling), C.POINTER(C.c_double), _P, _P, _I64,
                                    C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_int, StaticBeams, _P]),
    "bh_tr_params": (C.c_int, [C.POINTER(Tiling), C.POINTER(C.c_double), _P, _P, _I64, C.c_int,
                               C.c_double, C.c_double, C.c_double, C.c_int, StaticBeams, _P]),
    "bh_mswpt_combine_velocity": (C.c_int, [C.POINTER(Tiling), C.POINTER(C.c_double), C.c_double,
                                            _P, _P, C.c_int, _P, _P, _P]),
    "bh_forward_sum_modes": (C.c_int, [C.POINTER(FwdBeams), C.c_int, _I64, C.c_int, C.c_double, _P,
                                       _I64,
                                       C.POINTER(Segs), C.c_double, C.c_int, C.c_int, _P, _P, _P,
                                       _SZ, _P]),
    "bh_forward_sum_culled_workspace_size": (_SZ, [_I64, _I64]),
    "bh_forward_culled_workspace_size": (_SZ, [_I64, C.c_int, _I64, _I64]),
    "bh_forward_sum_culled": (C.c_int, [C.POINTER(FwdBeams), C.c_int, C.c_int, C.c_double, _P,
                                        _I64, C.POINTER(Segs), C.c_double, C.c_int, _P, _P, _P,
                                        _SZ, _P]),
    "bh_static_sum_culled_workspace_size": (_SZ, [_I64, C.c_int, _I64]),
    "bh_static_sum_culled": (C.c_int, [C.POINTER(StaticBeams), C.c_int, C.POINTER(Segs),
                                       C.c_double, C.c_double, C.c_int, _P, _P, _P, _SZ, _P]),
    "bh_forward_transpose_workspace_size": (_SZ, [_I64, _I64, _I64]),
    "bh_forward_transpose": (C.c_int, [C.POINTER(FwdBeams), C.c_int, C.c_int, C.c_double, _P, _I64,
                                       C.POINTER(Segs), C.c_double, C.c_int, _P, C.c_int, _P, _P,
                                       _SZ, _P]),
    "bh_coeff_scatter": (C.c_int, [_P, _P, _I64, C.c_double, _I64, _P, _P]),
    "bh_coeff_gather": (C.c_int, [_P, _I64, _P, _I64, _P, _P]),
    "bh_mswpt_adjoint": (C.c_int, [C.POINTER(Tiling), _P, _P, _P, C.c_int, _P, _P, _SZ, _P]),
    "bh_speed_eval": (C.c_int, [C.POINTER(Speed), _P, _I64, _P, _P, _P, _P]),
    "bh_ode_forward": (C.c_int, [C.POINTER(FwdBeams), C.c_int, C.c_int, C.POINTER(Speed), _P, _I64,
                                 C.c_double, C.c_double, _I64, _P, _P, _P]),
    "bh_boundary_terminal": (C.c_int, [C.POINTER(Tiling), C.POINTER(C.c_double), _P, _P, _I64,
                                       C.c_int, C.c_double, C.c_double, C.POINTER(Speed),
                                       C.c_double, C.c_int, C.c_int, StaticBeams, _P, _P, _P]),
    "bh_ode_static": (C.c_int, [C.POINTER(StaticBeams), C.c_int, C.POINTER(Speed), _P, _P, _P,
                                C.c_double, C.c_double, C.c_double, C.c_double, _I64, _P, _P]),
    "bh_forward_sum_traj": (C.c_int, [C.POINTER(FwdBeams), C.c_int, C.c_int, _P, _P, _I64,
                                      C.POINTER(Segs), C.c_double, C.c_int, _P, _P, _P, _SZ, _P]),
    "bh_forward_windows_workspace_size": (_SZ, [_I64, C.c_int]),
    "bh_forward_batch_windows": (C.c_int, [C.POINTER(BatchSolve), C.c_int, C.c_int, C.c_double,
                                           _P]),
    "bh_forward_batch_plan_size": (_SZ, [C.c_int]),
    "bh_forward_batch_plan": (C.c_int, [C.POINTER(BatchSolve), C.c_int, C.c_int, C.c_double, C.c_int,
                                        _P, _SZ, C.POINTER(C.c_int64), _P]),
    "bh_forward_batch_run": (C.c_int, [_P, C.c_int, _I64, C.c_int, C.c_int, C.c_int, _P, _P]),
    "bh_nn_gemm_workspace_size": (_SZ, [_I64, _I64, _I64, C.c_int]),
    "bh_nn_gemm": (C.c_int, [_P, C.c_int, _I64, _P, C.c_int, _I64, _I64, _I64, _I64, C.c_float,
                             _P, C.c_int, _P, C.c_int, _I64, C.c_int, _P, _SZ, _P]),
    "bh_nn_conv3": (C.c_int, [_P, _I64, C.c_int, C.c_int, C.c_int, _P, C.c_int, _P, C.c_int,
                              C.c_float, _P, C.c_int, _P]),
    "bh_nn_conv3_wflip": (C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    "bh_nn_conv3_dw_workspace_size": (_SZ, [_I64, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bh_nn_conv3_dw": (C.c_int, [_P, _P, _I64, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _SZ,
                                 _P]),
    "bh_nn_pool2": (C.c_int, [_P, _I64, C.c_int, C.c_int, C.c_int, _P, C.c_int, _P]),
    "bh_nn_up2": (C.c_int, [_P, _I64, C.c_int, C.c_int, C.c_int, _P, C.c_int, _P]),
    "bh_nn_add": (C.c_int, [_P, _P, _I64, _P, _P]),
    "bh_nn_relu_bwd": (C.c_int, [_P, _P, _I64, _P, _P]),
    "bh_nn_colsum_workspace_size": (_SZ, [_I64, C.c_int]),
    "bh_nn_colsum": (C.c_int, [_P, _I64, C.c_int, _P, _P, _SZ, _P]),
    "bh_nn_mse_workspace_size": (_SZ, [_I64]),
    "bh_nn_mse": (C.c_int, [_P, C.c_int, _P, _I64, _P, C.c_int, _P, _SZ, _P]),
    "bh_nn_adam": (C.c_int, [_P, _P, _P, _P, _I64, C.c_float, C.c_float, C.c_float, C.c_float,
                             _I64, _P, _P]),
    "bh_nn_cast": (C.c_int, [_P, C.c_int, _I64, C.c_float, _P, _P]),
    "bh_static_sum_workspace_size": (_SZ, [_I64, C.c_int]),
    "bh_static_sum": (C.c_int, [C.POINTER(StaticBeams), C.c_int, _P, _I64, C.c_double, C.c_int,
                                _P, _P, _SZ, _P]),
}

_lib = None


class BeamaxHIPError(RuntimeError):
    """A C-ABI call returned a non-zero status."""


def load():
    """Load the HIP library (once). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m beamax_amd.build` "
            "(the MSGB operator has no CPU fallback)."
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


# Tensors whose raw addresses were taken by ``ptr`` since the last ``call``
# returned, per thread.  A pointer argument written as ptr(make_tensor(...))
# would otherwise let the temporary die before the call is made, and the
# caching allocator could hand its memory to the next argument's temporary.
_held = threading.local()


def call(name, *args):
    """Call C-ABI function `name`; raise BeamaxHIPError on a non-zero status.

    Every tensor passed through ``ptr`` for this call stays referenced until
    the C function has returned (its kernels are then enqueued, and the
    stream orders any later reuse of the memory after them)."""
    lib = load()
    try:
        rc = getattr(lib, name)(*args)
    finally:
        _held.tensors = []
    if rc != 0:
        msg = lib.bh_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise BeamaxHIPError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t):
    """Raw device pointer of a torch tensor (None -> NULL); the tensor is held
    until the next ``call`` on this thread returns."""
    if t is None:
        return None
    held = getattr(_held, "tensors", None)
    if held is None:
        held = _held.tensors = []
    held.append(t)
    return C.c_void_p(t.data_ptr())


STATUS_BITS = {1: "forward parameter kernel met a beam index outside [-1, total)",
               2: "adjoint / time-reversal parameter kernel met a beam index outside [-1, total)",
               4: "coefficient scatter/gather met an index outside [-1, total)"}


def device_status(clear: bool = True) -> int:
    """The index-guard status bits of the current device (synchronises the
    current stream); 0 when no kernel met an out-of-range index."""
    bits = C.c_uint32(0)
    call("bh_device_status", int(clear), C.byref(bits), stream_handle())
    return int(bits.value)


def check_device_status() -> None:
    """Raise BeamaxHIPError if any index guard fired since the last check."""
    bits = device_status(clear=True)
    if bits:
        raise BeamaxHIPError("; ".join(m for b, m in STATUS_BITS.items() if bits & b))


def stream_handle():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def dbl_array(vals):
    arr = (C.c_double * max(1, len(vals)))()
    for i, v in enumerate(vals):
        arr[i] = float(v)
    return arr


def int32_array(vals):
    arr = (C.c_int32 * max(1, len(vals)))()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


def workspace(nbytes, device):
    import torch
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def as_f64_cols(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))

"""bh_topn_noalloc (the top_n selection on the library's fixed device
scratch array) against bh_topn with a caller workspace: two calls in a row,
the second on a smaller input, bit-exact indices and values; sizes beyond
the array are rejected before any GPU work."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _topn_pair(gpu, c, k):
    import torch
    from beamax_amd import _lib
    n = c.numel()
    kk = min(k, n)
    ref_i = torch.empty(kk, dtype=torch.int64, device=gpu)
    ref_v = torch.empty(kk, dtype=torch.complex128, device=gpu)
    ws = _lib.workspace(_lib.load().bh_topn_workspace_size(n, kk), gpu)
    _lib.call("bh_topn", _lib.ptr(c), n, kk, _lib.ptr(ref_i), _lib.ptr(ref_v), _lib.ptr(ws),
              ws.numel(), _lib.stream_handle())
    got_i = torch.full((kk,), -7, dtype=torch.int64, device=gpu)
    got_v = torch.full((kk,), np.nan, dtype=torch.complex128, device=gpu)
    _lib.call("bh_topn_noalloc", _lib.ptr(c), n, kk, _lib.ptr(got_i), _lib.ptr(got_v),
              _lib.stream_handle())
    torch.cuda.synchronize()
    return ref_i.cpu().numpy(), ref_v.cpu().numpy(), got_i.cpu().numpy(), got_v.cpu().numpy()


def test_noalloc_topn_matches_topn_on_consecutive_calls(gpu):
    import torch
    from oracle import topn as otopn
    rng = np.random.default_rng(2024)
    # the adjoint face size of the 512^2 workload, then a smaller tie-heavy one
    cases = [(2 * 1024 * 512, 65536, False), (40000, 3000, True)]
    for n, k, ties in cases:
        if ties:
            c = (rng.integers(0, 5, n) + 1j * rng.integers(0, 3, n)).astype(np.complex128)
        else:
            c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        cd = torch.as_tensor(c, device=gpu)
        ri, rv, gi, gv = _topn_pair(gpu, cd, k)
        assert np.array_equal(ri, gi)
        assert np.array_equal(rv.view(np.float64), gv.view(np.float64))
        oi, ov = otopn.top_n(c, k)
        assert np.array_equal(oi, gi)


def test_noalloc_topn_rejects_sizes_beyond_its_scratch(gpu):
    import torch
    from beamax_amd import _lib
    c = torch.zeros(1, dtype=torch.complex128, device=gpu)
    out_i = torch.empty(1, dtype=torch.int64, device=gpu)
    out_v = torch.empty(1, dtype=torch.complex128, device=gpu)
    # sizes are checked on the host before any pointer is used
    with pytest.raises(_lib.BeamaxHIPError, match="exceed its fixed scratch"):
        _lib.call("bh_topn_noalloc", _lib.ptr(c), (1 << 22) + 1, 10, _lib.ptr(out_i),
                  _lib.ptr(out_v), _lib.stream_handle())
    with pytest.raises(_lib.BeamaxHIPError, match="exceed its fixed scratch"):
        _lib.call("bh_topn_noalloc", _lib.ptr(c), 1 << 20, (1 << 18) + 1, _lib.ptr(out_i),
                  _lib.ptr(out_v), _lib.stream_handle())

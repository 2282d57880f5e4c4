"""The benchmark workloads at their full sizes on the GPU.

* 512^2 forward + four-face adjoint (4 levels, top_n 65536, 2044 sensors,
  2414 steps; faces cropped to 1024 steps, top_n 65536 each), FP32-split
  with the timed 2^-32 envelope cut, against the dense FP64 kernels on the
  same problem (rel-L2 <= 1e-5), and the dense FP64 kernels against the
  oracle: every beam on a deterministic 1/128 subsample of the (time,
  sensor) grid (every 16th step x every 8th sensor) and of the image (every
  128th pixel), with the oracle's own MSWPT, top_n, beam parameters and the
  four-face composition of ``oracle.adjoint.four_edge_adjoint``.
* 128^3 forward + adjoint (top_n 262144, 16384 sensors on the axis-0 face,
  740 steps; the face data cropped to 256 steps), FP32-split at 2^-32
  against dense FP64 sums of the same beams on every 16th sensor and every
  64th voxel.
* the 64-solve batch (grids 128^2-512^2, 1k-64k beams, 400-2000 steps) as
  one FP32-split launch at 2^-32 against each solve's dense FP64 sum.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _n(x):
    return x.detach().cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (np.linalg.norm(a) + 1e-300))


def max_rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(a)))


def test_512_forward_adjoint_full_size(gpu):
    import torch
    import bench
    from adjoint_cases import cpp_field_sum
    from oracle import adjoint as oadj
    from oracle import beams as obeams
    from oracle import cpu as ocpu
    from oracle import topn as otopn
    from oracle.decomposition import Tiling
    from oracle.transforms import MSWPT as OMSWPT
    P = bench.build_problem(gpu, 32.0)
    d32, i32, _, _ = bench.run_step(P, "fp32_split")
    d64, i64, _, _ = bench.run_step(P, "fp64")
    torch.cuda.synchronize()
    d32, i32 = _n(d32).astype(np.float64), _n(i32).astype(np.float64)
    d64, i64 = _n(d64), _n(i64)
    assert d64.shape == (2414, 2044) and i64.size == 512 * 512
    assert rel(d64, d32) < 1e-5 and max_rel(d64, d32) < 1e-5
    assert rel(i64, i32) < 1e-5 and max_rel(i64, i32) < 1e-5
    # forward: the oracle's own pipeline, every beam, on the subsample
    dom, N = P["dom"], (512, 512)
    p0 = _n(P["p0"])
    ow = OMSWPT(Tiling(4, N, (4, 8, 16, 32), (1, 1)), 2, "rectangular_mirror")
    L = dom.grid_size
    cpos = obeams.compute_coefficients_pos_only(ow, p0, 1500.0, L)
    ridx, rval = otopn.top_n(cpos, 65536)
    _, gidx, _ = P["solvers"]["fp64"].prepare_forward_beams(P["p0"], dom, P["wpt"])
    assert np.array_equal(np.sort(ridx), np.sort(_n(gidx)))
    p, al, x0, om, a0, mode = obeams.forward_parameters(ow, ridx, L, N)
    dup = lambda v: np.concatenate([v, v])
    ts, sens = P["ts"], P["sensors"].positions
    ref = ocpu.forward_sum(dup(x0), dup(p), dup(al), dup(a0 * rval), dup(om),
                           np.concatenate([mode, -mode]), 1500.0, ts[::16], sens[::8])
    assert rel(ref, d64[::16, ::8]) < 1e-12
    assert rel(ref, d32[::16, ::8]) < 1e-5
    # adjoint: the oracle's four-face composition on every 128th pixel,
    # first with its own top_n per face (the selection must agree with the
    # GPU's up to one conjugate pair at rank N), then on the GPU's selection
    pix = np.arange(0, 512 * 512, 128)
    sel = [_n(b.idx) for b in P["adj"]["fp64"].face_beams]
    _, own = oadj.four_edge_adjoint(d64, N, dom.dx, 1500.0, ts, 1024, 65536,
                                    field_sum=cpp_field_sum, pixels=pix[:8])
    for o, g in zip(own, sel):
        assert len(set(o[0].tolist()) ^ set(g.tolist())) <= 2
    qref, _ = oadj.four_edge_adjoint(d64, N, dom.dx, 1500.0, ts, 1024, 65536, selections=sel,
                                     field_sum=cpp_field_sum, pixels=pix)
    assert rel(qref, i64.ravel()[pix]) < 1e-11
    assert rel(qref, i32.ravel()[pix]) < 1e-5


def test_128cubed_forward_adjoint_against_dense_fp64(gpu):
    import torch
    import bench
    import beamax_amd as bx
    from beamax_amd.gb import gb_solvers
    P = bench.build_problem_3d(gpu, 32.0)
    d32, i32, beams, ab = bench.run_step_3d(P, "fp32_split")
    dense = bx.MSGBSolver(262144, "top_n", 128, "spatial", gb_solvers.solve_hom_diag,
                          "scan_real", precision="fp64", cull_log2=0.0)
    sens = P["sensors"].positions[::16]
    ref_fwd = _n(dense.sum_forward(beams, P["dom"], sens, P["ts"]))
    vox = P["points"][::64]
    ref_adj = _n(dense.sum_static(ab, vox, 1.0 / 1500.0 ** 2))
    torch.cuda.synchronize()
    got_fwd = _n(d32).astype(np.float64)[:, ::16]
    got_adj = _n(i32).astype(np.float64).ravel()[::64]
    assert ref_fwd.shape == (740, 1024) and ref_adj.size == 128 ** 3 // 64
    assert rel(ref_fwd, got_fwd) < 1e-5 and max_rel(ref_fwd, got_fwd) < 1e-5
    assert rel(ref_adj, got_adj) < 1e-5 and max_rel(ref_adj, got_adj) < 1e-5


def test_64_solve_batch_against_dense_fp64(gpu):
    import torch
    import beamax_amd as bx
    from beamax_amd.batch import BatchedForward
    from beamax_amd.gb import gb_solvers
    from beamax_amd.workloads import batch_forward_problems
    probs = batch_forward_problems(gpu, n_solves=64, seed=0)
    s = bx.MSGBSolver(65536, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                      precision="fp32_split", cull_log2=32.0)
    bf = BatchedForward(s, probs)
    outs = [o.clone() for o in bf.run()]
    dense = bx.MSGBSolver(65536, "top_n", 128, "spatial", gb_solvers.solve_hom_diag,
                          "scan_real", precision="fp64", cull_log2=0.0)
    sizes = set()
    for p, it, o in zip(probs, bf.items, outs):
        ref = _n(dense.sum_forward(it["beams"], p.domain, p.points, p.ts))
        got = _n(o).astype(np.float64)
        assert np.all(np.isfinite(got))
        assert rel(ref, got) < 1e-5, (p.domain.N, it["k"], p.ts.size)
        sizes.add(p.domain.N[0])
    torch.cuda.synchronize()
    assert len(probs) == 64 and sizes == {128, 192, 256, 384, 512}

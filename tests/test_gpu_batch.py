"""Batched forward solves: one summation launch over many solves of
different grids, beam counts and time grids, its HIP-graph capture and the
tile autotuner, against the single-solve culled sums."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _n(x):
    return x.detach().cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (np.linalg.norm(a) + 1e-300))


def _batch(gpu, prec, n=6, seed=5):
    import beamax_amd as bx
    from beamax_amd.batch import BatchedForward
    from beamax_amd.gb import gb_solvers
    from beamax_amd.workloads import batch_forward_problems
    probs = batch_forward_problems(gpu, n_solves=n, seed=seed, sizes=(128, 192, 256),
                                   nt_range=(60, 200), beam_range=(200, 3000))
    s = bx.MSGBSolver(1000, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                      precision=prec, cull_log2=40.0)
    return s, probs, BatchedForward(s, probs)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32_split", 1e-6)])
def test_batched_sum_matches_single_solves(gpu, prec, tol):
    import torch
    s, probs, bf = _batch(gpu, prec)
    outs = [o.clone() for o in bf.run()]
    torch.cuda.synchronize()
    assert bf.dense_evals() > 0
    for p, it, o in zip(probs, bf.items, outs):
        one = s.sum_forward(it["beams"], p.domain, p.points, p.ts)
        got = _n(o).astype(np.float64)
        assert np.all(np.isfinite(got))
        assert rel(_n(one).astype(np.float64), got) < tol
        # and the whole single-solve forward (same top_n selection)
        s1 = type(s)(it["k"], "top_n", 128, "spatial", s.ode_solver, "scan_real",
                     precision=prec, cull_log2=40.0)
        ref = _n(s1.forward(p.p0, p.domain, p.points, p.ts, p.wpt)).astype(np.float64)
        assert rel(ref, got) < tol


def test_batched_graph_replay_and_tiles(gpu):
    import torch
    s, probs, bf = _batch(gpu, "fp32_split", n=5, seed=6)
    eager = [o.clone() for o in bf.run()]
    bf.capture()
    for rep in range(3):
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device=gpu).sum()   # eager work between replays
        out = bf.replay()
        torch.cuda.synchronize()
        for a, b in zip(eager, out):
            assert torch.equal(a, b), rep
        del junk
    # the graph reads the inputs it captured: doubling every p0 doubles the
    # data (same selection, amplitudes scaled by two; the FP32 terms carry
    # log2 of the amplitude in their exponent, so to rounding, not bit for bit)
    for p in probs:
        p.p0.mul_(2.0)
    out = bf.replay()
    torch.cuda.synchronize()
    for a, b in zip(eager, out):
        assert rel(_n(2.0 * a).astype(np.float64), _n(b).astype(np.float64)) < 1e-6
    for p in probs:
        p.p0.mul_(0.5)
    # every tile shape sums each point in the same order: identical results
    best, times = bf.tune(reps=1)
    assert best in (8, 16, 32) and len(times) == 3
    for gpb in (8, 32):
        bf.set_tile(gpb)
        out = bf.run()
        torch.cuda.synchronize()
        for a, b in zip(eager, out):
            assert torch.equal(a, b), gpb

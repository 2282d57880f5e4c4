"""The whole forward+adjoint step (forward MSWPT, top_n and beam sums; the
four faces' data MSWPT, top_n and records on their own streams; the image
sum) captured as one HIP graph: every replay reproduces the eager step bit
for bit, with eager work between the replays, and the graph reads the input
it captured."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _n(x):
    return x.detach().cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (np.linalg.norm(a) + 1e-300))


def test_forward_adjoint_step_graph_replay(gpu):
    import torch
    import beamax_amd as bx
    from beamax_amd.gb import gb_solvers
    from beamax_amd.solvers.faces import FourEdgeAdjoint, four_edge_mask
    from beamax_amd.workloads import packets_2d

    N = (128, 128)
    dx = (5e-5, 5e-5)
    dom = bx.Domain(N=N, dx=dx, c=1500.0, cfl=0.3, periodic=(False, False))
    wpt = bx.MSWPT(bx.DyadicDecomposition(3, N, (4, 8, 16), (1, 1)), 2, "rectangular_mirror",
                   device=gpu)
    sensors = bx.Sensor(dom, binary_mask=four_edge_mask(N))
    ts = dom.generate_time_domain()
    p0 = torch.as_tensor(packets_2d(N, dx, 16, seed=3), device=gpu)
    s = bx.MSGBSolver(4096, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                      precision="fp32_split", cull_log2=32.0)
    fa = FourEdgeAdjoint(s, dom, ts, 256, device=gpu)

    def step():
        beams, _, _ = s.prepare_forward_beams(p0, dom, wpt)
        data = s.sum_forward(beams, dom, sensors.positions, ts)
        ab = fa.beams(data)
        img = s.sum_static(ab, fa.points, 1.0 / dom.c ** 2, dx=dom.dx)
        return data, img

    for _ in range(2):                      # warm every workspace and cache
        data, img = step()
    data, img = data.clone(), img.clone()
    torch.cuda.synchronize()
    assert float(data.abs().max()) > 0 and float(img.abs().max()) > 0
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        gdata, gimg = step()
    for rep in range(3):
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device=gpu).sum()   # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(gdata, data), rep
        assert torch.equal(gimg, img), rep
        del junk
    # the graph reads the p0 it captured: doubling it doubles the data and the
    # image (same selections, amplitudes scaled by two; to rounding)
    p0.mul_(2.0)
    g.replay()
    torch.cuda.synchronize()
    assert rel(_n(2.0 * data).astype(np.float64), _n(gdata).astype(np.float64)) < 1e-6
    assert rel(_n(2.0 * img).astype(np.float64), _n(gimg).astype(np.float64)) < 1e-5
    p0.mul_(0.5)
    del g

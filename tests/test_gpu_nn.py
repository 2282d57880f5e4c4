"""The learned reconstruction network's kernels on the GPU against a plain
PyTorch FP32 (CPU) reference of the same ops: the bf16 MFMA GEMM in every
operand layout (with bias/ReLU epilogue and split-K), the convolution
forward and both gradients, and the whole U-Net forward/backward against a
CPU twin with the same weights; and a training smoke test."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf(t):
    import torch
    return t.to(torch.bfloat16).to(torch.float32)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-300))


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_layouts_match_torch(gpu, ta, tb):
    import torch
    from beamax_amd import _lib
    g = torch.Generator().manual_seed(10 + 2 * ta + tb)
    M, N, K = 104, 72, 136
    A = _bf(torch.randn(M, K, generator=g))
    B = _bf(torch.randn(K, N, generator=g))
    bias = torch.randn(N, generator=g)
    ref = A.double() @ B.double()
    As = (A.t() if ta else A).contiguous().to(torch.bfloat16).to(gpu)
    Bs = (B.t() if tb else B).contiguous().to(torch.bfloat16).to(gpu)
    lda = M if ta else K
    ldb = K if tb else N
    C = torch.empty(M, N, dtype=torch.float32, device=gpu)
    _lib.call("bh_nn_gemm", _lib.ptr(As), ta, lda, _lib.ptr(Bs), tb, ldb, M, N, K, 1.0, None, 0,
              _lib.ptr(C), 0, N, 1, None, 0, _lib.stream_handle())
    assert rel(C.cpu().numpy(), ref.numpy()) < 1e-6
    # epilogue: 0.5 * AB + bias, ReLU, bf16 output
    bd = bias.to(gpu)
    Cb = torch.empty(M, N, dtype=torch.bfloat16, device=gpu)
    _lib.call("bh_nn_gemm", _lib.ptr(As), ta, lda, _lib.ptr(Bs), tb, ldb, M, N, K, 0.5,
              _lib.ptr(bd), 1, _lib.ptr(Cb), 1, N, 1, None, 0, _lib.stream_handle())
    r2 = torch.relu(0.5 * ref + bias.double())
    got = Cb.float().cpu().double()
    assert float((got - r2).abs().max()) <= 2 ** -7 * float(r2.abs().max()) + 1e-6
    # split-K over a longer K
    K2 = 1000
    A2 = _bf(torch.randn(M, K2, generator=g))
    B2 = _bf(torch.randn(K2, N, generator=g))
    A2s = (A2.t() if ta else A2).contiguous().to(torch.bfloat16).to(gpu)
    B2s = (B2.t() if tb else B2).contiguous().to(torch.bfloat16).to(gpu)
    lib = _lib.load()
    nb = lib.bh_nn_gemm_workspace_size(M, N, K2, 5)
    ws = torch.empty(nb // 4, dtype=torch.float32, device=gpu)
    C2 = torch.empty(M, N, dtype=torch.float32, device=gpu)
    _lib.call("bh_nn_gemm", _lib.ptr(A2s), ta, M if ta else K2, _lib.ptr(B2s), tb,
              K2 if tb else N, M, N, K2, 1.0, None, 0, _lib.ptr(C2), 0, N, 5, _lib.ptr(ws), nb,
              _lib.stream_handle())
    assert rel(C2.cpu().numpy(), (A2.double() @ B2.double()).numpy()) < 1e-6


def test_conv_forward_and_gradients_match_torch(gpu):
    """One 3x3 conv layer through the implicit-GEMM kernel (forward), and
    its weight, bias and input gradients (implicit GEMM over the pixels,
    column sums, the convolution with flipped weights) against
    torch.nn.functional.conv2d and autograd in FP32 on the CPU."""
    import torch
    import torch.nn.functional as Fn
    from beamax_amd.learned import UNet
    net = UNet((8, 16, 16, 16), device=gpu, seed=3)
    conv = net.layers["e2a"]            # 8 -> 16 channels, ReLU
    g = torch.Generator().manual_seed(4)
    B, H, W = 2, 12, 20
    x = _bf(torch.randn(B, H, W, conv.cin, generator=g))
    w = net.params.view(conv.wi, (conv.cout, 9, conv.cin)).cpu()
    net.params.view(conv.bi, (conv.cout,)).copy_(torch.randn(conv.cout, generator=g).to(gpu))
    b = net.params.view(conv.bi, (conv.cout,)).cpu()
    wb = _bf(w)
    # torch reference: NCHW, weight (cout, cin, 3, 3) with tap = 3 (dy+1) + (dx+1)
    xt = x.permute(0, 3, 1, 2).clone().requires_grad_(True)
    wt = wb.reshape(conv.cout, 3, 3, conv.cin).permute(0, 3, 1, 2).clone().requires_grad_(True)
    bt = b.clone().requires_grad_(True)
    z = Fn.conv2d(xt, wt, bt, padding=1)
    y = torch.relu(z)
    yd = net._conv_fwd(conv, x.to(torch.bfloat16).to(gpu), B, H, W)
    yref = y.permute(0, 2, 3, 1)
    assert rel(yd.float().cpu().numpy(), yref.detach().numpy()) < 5e-3
    dy = _bf(torch.randn(B, H, W, conv.cout, generator=g))
    # backward through the same ReLU mask as the device (its own bf16 output)
    mask = (yd.float().cpu() > 0).float()
    (z * (dy * mask).permute(0, 3, 1, 2)).sum().backward()
    dx = net._conv_bwd(conv, x.to(torch.bfloat16).to(gpu), yd, dy.to(torch.bfloat16).to(gpu),
                       B, H, W)
    dW = net.params.view(conv.wi, (conv.cout, 9, conv.cin), grad=True).cpu()
    db = net.params.view(conv.bi, (conv.cout,), grad=True).cpu()
    dW_ref = wt.grad.permute(0, 2, 3, 1).reshape(conv.cout, 9, conv.cin)
    assert rel(dW.numpy(), dW_ref.numpy()) < 1e-5
    assert rel(db.numpy(), bt.grad.numpy()) < 1e-5
    assert rel(dx.float().cpu().numpy(), xt.grad.permute(0, 2, 3, 1).numpy()) < 5e-3


@pytest.mark.parametrize("cin,cout,shape", [
    (8, 24, (2, 10, 12)), (24, 48, (1, 17, 9)), (48, 96, (3, 8, 8)), (96, 192, (2, 6, 10)),
    (192, 96, (1, 5, 7)), (16, 8, (2, 64, 40)), (64, 64, (1, 3, 3)), (24, 24, (4, 48, 64)),
    (8, 24, (2, 16, 24)), (48, 64, (2, 24, 32)), (24, 8, (3, 8, 16)), (48, 48, (1, 40, 8)),
    (48, 24, (2, 32, 32)), (24, 16, (3, 7, 24)), (48, 32, (2, 1, 16)),
    (96, 80, (2, 9, 16)), (72, 192, (1, 6, 8)), (192, 96, (2, 12, 24)), (8, 16, (3, 5, 16))])
def test_implicit_conv_kernels_match_torch(gpu, cin, cout, shape):
    """The three convolution entry points on their own, over every output
    tile width (8 to 192 channels, several k and channel tiles), images
    smaller than a tile, widths that are not multiples of 8, heights that are
    not multiples of the weight gradient's 4-row tiles, and pixel counts
    that split the weight gradient into many chunks (both weight-gradient
    kernels: the whole-tile one for 8 input channels or a multiple of 24 --
    slices of 8, 24 or 48, output tiles of 48 or 64 -- at widths that are
    multiples of 8, the k-tiled one otherwise; both
    forward kernels: halo tiles and gathered operands): forward (bias, ReLU,
    alpha, bf16 and FP32 output), input gradient through the flipped weights,
    weight gradient, all against FP32 torch on the CPU."""
    import torch
    import torch.nn.functional as Fn
    from beamax_amd import _lib
    B, H, W = shape
    g = torch.Generator().manual_seed(cin * 1000 + cout)
    x = _bf(torch.randn(B, H, W, cin, generator=g))
    w = _bf(torch.randn(cout, 9, cin, generator=g) / (3.0 * cin ** 0.5))
    bias = torch.randn(cout, generator=g)
    dz = _bf(torch.randn(B, H, W, cout, generator=g))
    wt = w.reshape(cout, 3, 3, cin).permute(0, 3, 1, 2).double().requires_grad_(True)
    xt = x.permute(0, 3, 1, 2).double().requires_grad_(True)
    z = Fn.conv2d(xt, wt, None, padding=1)
    (z * dz.permute(0, 3, 1, 2).double()).sum().backward()
    zref = z.detach().permute(0, 2, 3, 1)
    st = _lib.stream_handle()
    xd = x.to(torch.bfloat16).to(gpu)
    wd = w.reshape(cout, 9 * cin).to(torch.bfloat16).to(gpu)
    bd = bias.to(gpu)
    y = torch.empty(B, H, W, cout, dtype=torch.float32, device=gpu)
    _lib.call("bh_nn_conv3", _lib.ptr(xd), B, H, W, cin, _lib.ptr(wd), cout, None, 0, 1.0,
              _lib.ptr(y), 0, st)
    assert rel(y.cpu().numpy(), zref.numpy()) < 1e-5
    yb = torch.empty(B, H, W, cout, dtype=torch.bfloat16, device=gpu)
    _lib.call("bh_nn_conv3", _lib.ptr(xd), B, H, W, cin, _lib.ptr(wd), cout, _lib.ptr(bd), 1, 0.5,
              _lib.ptr(yb), 1, st)
    r2 = torch.relu(0.5 * zref + bias.double())
    assert float((yb.float().cpu().double() - r2).abs().max()) <= \
        2 ** -7 * float(r2.abs().max()) + 1e-6
    dzd = dz.to(torch.bfloat16).to(gpu)
    wf = torch.empty(cin * 9 * cout, dtype=torch.bfloat16, device=gpu)
    _lib.call("bh_nn_conv3_wflip", _lib.ptr(wd), cout, cin, _lib.ptr(wf), st)
    dx = torch.empty(B, H, W, cin, dtype=torch.float32, device=gpu)
    _lib.call("bh_nn_conv3", _lib.ptr(dzd), B, H, W, cout, _lib.ptr(wf), cin, None, 0, 1.0,
              _lib.ptr(dx), 0, st)
    assert rel(dx.cpu().numpy(), xt.grad.permute(0, 2, 3, 1).numpy()) < 1e-5
    lib = _lib.load()
    nb = lib.bh_nn_conv3_dw_workspace_size(B, H, W, cin, cout)
    ws = torch.empty(max(nb // 4, 1), dtype=torch.float32, device=gpu)
    dW = torch.full((cout, 9 * cin), float("nan"), dtype=torch.float32, device=gpu)
    _lib.call("bh_nn_conv3_dw", _lib.ptr(xd), _lib.ptr(dzd), B, H, W, cin, cout, _lib.ptr(dW),
              _lib.ptr(ws), nb, st)
    dW_ref = wt.grad.permute(0, 2, 3, 1).reshape(cout, 9 * cin)
    assert rel(dW.cpu().numpy(), dW_ref.numpy()) < 1e-5


@pytest.mark.parametrize("C,shape", [(8, (2, 6, 10)), (24, (3, 4, 8)), (5, (2, 6, 4))])
def test_elementwise_kernels_bit_exact(gpu, C, shape):
    """Pooling, upsampling, their adjoints, the skip add and the ReLU
    backward, bit for bit against the same FP32 operations on the CPU (2x2
    sums in (dy, dx) order, round to nearest even), both on 16-byte aligned
    operands with C a multiple of 8 (the 8-channel chunk kernels) and on
    operands one element off alignment (the per-element kernels); the column
    sums of the bias gradient on both paths against an FP64 sum."""
    import torch
    from beamax_amd import _lib
    B, H, W = shape
    g = torch.Generator().manual_seed(C * 7 + H)
    big = torch.randn(B, H, W, C, generator=g).to(torch.bfloat16)
    small = torch.randn(B, H // 2, W // 2, C, generator=g).to(torch.bfloat16)
    other = torch.randn(B, H, W, C, generator=g).to(torch.bfloat16)
    f = big.float()
    quad = ((f[:, 0::2, 0::2] + f[:, 0::2, 1::2]) + f[:, 1::2, 0::2]) + f[:, 1::2, 1::2]
    rep = lambda t: t.repeat_interleave(2, 1).repeat_interleave(2, 2)
    want = {
        ("pool2", 0): (big, (0.25 * quad).to(torch.bfloat16)),
        ("pool2", 1): (small, rep((0.25 * small.float()).to(torch.bfloat16))),
        ("up2", 0): (small, rep(small)),
        ("up2", 1): (big, quad.to(torch.bfloat16)),
    }
    st = _lib.stream_handle()

    def dev(t, off):
        buf = torch.zeros(t.numel() + 8, dtype=torch.bfloat16, device=gpu)
        v = buf[off:off + t.numel()]
        v.copy_(t.reshape(-1).to(gpu))
        return v

    bits = lambda t: t.reshape(-1).view(torch.int16).cpu()
    for off in (0, 1):
        for (name, bwd), (x, ref) in want.items():
            xd, yd = dev(x, off), dev(torch.zeros_like(ref), off)
            b, h, w = (B, H, W) if name == "pool2" else (B, H // 2, W // 2)
            _lib.call(f"bh_nn_{name}", _lib.ptr(xd), b, h, w, C, _lib.ptr(yd), bwd, st)
            assert torch.equal(bits(yd), bits(ref)), (name, bwd, off)
        n = big.numel()
        xd, od, zd = dev(big, off), dev(other, off), dev(torch.zeros_like(big), off)
        _lib.call("bh_nn_add", _lib.ptr(xd), _lib.ptr(od), n, _lib.ptr(zd), st)
        assert torch.equal(bits(zd), bits((f + other.float()).to(torch.bfloat16))), ("add", off)
        _lib.call("bh_nn_relu_bwd", _lib.ptr(od), _lib.ptr(xd), n, _lib.ptr(zd), st)
        relu_ref = torch.where(f > 0, other, torch.zeros_like(other))
        assert torch.equal(bits(zd), bits(relu_ref)), ("relu_bwd", off)
        # column sums (bias gradients) of big as [P][C]: FP32 partials
        P = n // C
        lib = _lib.load()
        ws = torch.empty(lib.bh_nn_colsum_workspace_size(P, C) // 4 + 1, dtype=torch.float32,
                         device=gpu)
        cs = torch.empty(C, dtype=torch.float32, device=gpu)
        _lib.call("bh_nn_colsum", _lib.ptr(xd), P, C, _lib.ptr(cs), _lib.ptr(ws),
                  ws.numel() * 4, st)
        assert rel(cs.cpu().numpy(), f.reshape(P, C).double().sum(0).numpy()) < 1e-6, off


class _Twin:
    """FP32 CPU twin of ``UNet`` (same weights, bf16-rounded as the device
    uses them; activations rounded to bf16 where the device stores them)."""

    def __init__(self, net):
        import torch
        self.net = net
        self.w = {}
        for name, conv in net.layers.items():
            W = _bf(net.params.view(conv.wi, (conv.cout, 9, conv.cin)).cpu())
            Wt = W.reshape(conv.cout, 3, 3, conv.cin).permute(0, 3, 1, 2).contiguous()
            self.w[name] = (Wt.requires_grad_(True),
                            net.params.view(conv.bi, (conv.cout,)).cpu().clone()
                            .requires_grad_(True))

    def conv(self, name, x):
        import torch
        import torch.nn.functional as Fn
        conv = self.net.layers[name]
        W, b = self.w[name]
        z = Fn.conv2d(x, W, b, padding=1)
        if conv.relu:
            z = torch.relu(z)
        if conv.out_f32:
            return z
        return z + (_bf(z) - z).detach()          # bf16 storage, straight-through gradient

    def forward(self, x):
        import torch
        import torch.nn.functional as Fn
        r = lambda t: t + (_bf(t) - t).detach()
        up = lambda t: Fn.interpolate(t, scale_factor=2, mode="nearest")
        pool = lambda t: r(Fn.avg_pool2d(t, 2))
        e1a = self.conv("e1a", x)
        s1 = self.conv("e1b", e1a)
        e2a = self.conv("e2a", pool(s1))
        s2 = self.conv("e2b", e2a)
        e3a = self.conv("e3a", pool(s2))
        s3 = self.conv("e3b", e3a)
        b = self.conv("b3", self.conv("b2", self.conv("b1", pool(s3))))
        d3 = self.conv("d3", r(self.conv("u3", up(b)) + s3))
        d2 = self.conv("d2", r(self.conv("u2", up(d3)) + s2))
        d1 = self.conv("d1", r(self.conv("u1", up(d2)) + s1))
        return self.conv("head", d1)[:, 0]


def test_unet_forward_backward_matches_cpu_twin(gpu):
    import torch
    from beamax_amd.learned import UNet
    net = UNet((8, 16, 24, 32), device=gpu, seed=5)
    g = torch.Generator().manual_seed(6)
    B, H, W = 2, 32, 48
    img = torch.randn(B, H, W, generator=g)
    tgt = torch.randn(B, H, W, generator=g)
    x = net.input_tensor(img.to(gpu))
    out = net.forward(x).float().cpu()
    loss, grad = net.loss_and_grad(tgt.to(gpu))
    net.backward(grad)
    twin = _Twin(net)
    xt = x.float().cpu().permute(0, 3, 1, 2).contiguous()
    ot = twin.forward(xt)
    assert rel(out.numpy(), ot.detach().numpy()) < 2e-2
    lt = ((ot - tgt) ** 2).mean()
    lt.backward()
    assert abs(float(loss) - float(lt.detach())) <= 2e-2 * float(lt.detach())
    for name, conv in net.layers.items():
        dW = net.params.view(conv.wi, (conv.cout, 9, conv.cin), grad=True).cpu()
        Wt, bt = twin.w[name]
        dW_ref = Wt.grad.permute(0, 2, 3, 1).reshape(conv.cout, 9, conv.cin)
        db = net.params.view(conv.bi, (conv.cout,), grad=True).cpu()
        assert rel(dW.numpy(), dW_ref.numpy()) < 6e-2, name
        assert rel(db.numpy(), bt.grad.numpy()) < 6e-2, name


def test_training_reduces_the_loss(gpu):
    import torch
    from beamax_amd.learned import UNet
    net = UNet((8, 16, 24, 32), device=gpu, seed=7)
    g = torch.Generator().manual_seed(8)
    B, H, W = 4, 32, 32
    yy, xx = torch.meshgrid(torch.arange(H), torch.arange(W), indexing="ij")
    tgt = torch.stack([torch.exp(-((yy - 8 - 4 * i) ** 2 + (xx - 16) ** 2) / 30.0)
                       for i in range(B)]).float()
    img = tgt + 0.1 * torch.randn(B, H, W, generator=g)
    losses = [float(net.step(img.to(gpu), tgt.to(gpu), 3e-3)) for _ in range(60)]
    print("loss", losses[0], "->", losses[-1], flush=True)
    assert losses[-1] < 0.2 * losses[0]


def test_learned_start_on_a_held_out_seed(gpu):
    """The committed network (trained on phantom seeds 1000-5095) gives CG on
    the normal equations a start that cuts the iterations to the 1e-3
    normal-equation residual on held-out seed 3."""
    from pathlib import Path
    import beamax_amd
    from beamax_amd.learned import UNet, evaluate, recon_problem
    w = Path(beamax_amd.__file__).parent / "weights" / "unet_recon256.safetensors"
    net = UNet.load(w, device=gpu)
    assert net.n_params > 1_000_000
    row = evaluate(net, recon_problem(gpu), [3], gpu)[0]
    print(row, flush=True)
    assert row["start_residual"] < 0.3
    assert row["learned_iters"] < row["plain_iters"]

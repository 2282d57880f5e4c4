"""Learned starting image for CG on the normal equations.

A small U-Net maps the four-face MSGB adjoint image of the data to an
estimate of p0; CG on the normal equations (``reconstruct.cgnr``) then starts
from that estimate (scaled by the least-squares factor on the data) instead
of from zero.  The network is trained on (adjoint image, p0) pairs that the
MSGB forward and adjoint operators of this package generate, and every
kernel of its forward and backward passes is the package's own HIP code
(``csrc/bh_nn.hip``): 3x3 convolutions as implicit bf16 MFMA GEMMs with FP32
accumulation (operands gathered straight from the activations), their
weight gradients as implicit GEMMs over the pixels (split into chunks added
in order) and their input gradients as the same convolution of the output
gradient with flipped weights; pooling, upsampling, skip additions, ReLU
masks, the loss and Adam.
The reference has no reconstruction solver, so there is nothing to mirror.

Architecture (``UNet``): NHWC bf16 activations; 4 levels with channels
(24, 48, 96, 192); two 3x3 conv + ReLU per level (three at the bottom),
2x2 average pooling down, nearest 2x upsampling + 3x3 conv + ReLU up,
additive skip connections, a 3x3 head with one output channel (FP32).
Inputs are divided by their RMS; the output is p0 / RMS(input) / ``kappa``.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib


def _ceil8(n: int) -> int:
    return (n + 7) // 8 * 8


def _call(name, *args):
    _lib.call(name, *args)


def _st():
    return _lib.stream_handle()


class _Params:
    """All weights in one flat FP32 buffer (with Adam moments, gradient and a
    bf16 copy); layers hold views."""

    def __init__(self, sizes, device):
        import torch
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        n = int(self.offsets[-1])
        # 16-byte alignment of every view: sizes are multiples of 8 elements
        assert all(int(s) % 8 == 0 for s in sizes)
        self.n = n
        self.w = torch.zeros(n, dtype=torch.float32, device=device)
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.wb = torch.zeros(n, dtype=torch.bfloat16, device=device)
        self.step = 0

    def view(self, i, shape, bf16=False, grad=False):
        a, b = int(self.offsets[i]), int(self.offsets[i + 1])
        src = self.wb if bf16 else (self.g if grad else self.w)
        return src[a:b].view(*shape)

    def refresh_bf16(self):
        _call("bh_nn_cast", _lib.ptr(self.w), 0, self.n, 1.0, _lib.ptr(self.wb), _st())

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _call("bh_nn_adam", _lib.ptr(self.w), _lib.ptr(self.g), _lib.ptr(self.m),
              _lib.ptr(self.v), self.n, float(lr), float(beta1), float(beta2), float(eps),
              self.step, _lib.ptr(self.wb), _st())


@dataclass
class Conv3:
    """3x3 same-padded convolution (+ bias, optional ReLU) on NHWC bf16."""
    cin: int            # input channels as stored (multiple of 8)
    cout: int           # output channels as stored (multiple of 8)
    relu: bool
    out_f32: bool = False
    wi: int = -1        # parameter slot of the weights [cout][9 cin]
    bi: int = -1        # parameter slot of the bias [cout]
    cout_live: int = 0  # channels that carry values (head: 1)

    def sizes(self):
        return [self.cout * 9 * self.cin, self.cout]


class _Work:
    """Scratch buffers grown on demand (ReLU-masked gradients, flipped
    weights, weight-gradient partials, loss partials)."""

    def __init__(self, device):
        self.device = device
        self.buf = {}

    def get(self, name, n, dtype):
        import torch
        t = self.buf.get(name)
        if t is None or t.numel() < n or t.dtype != dtype:
            t = torch.empty(max(n, 8), dtype=dtype, device=self.device)
            self.buf[name] = t
        return t[:n]


class UNet:
    """See the module docstring.  ``forward`` keeps what ``backward`` needs."""

    def __init__(self, channels=(24, 48, 96, 192), device=None, seed: int = 0):
        import torch
        self.device = torch.device("cuda") if device is None else torch.device(device)
        c1, c2, c3, c4 = channels
        self.channels = tuple(channels)
        L = {}
        L["e1a"] = Conv3(8, c1, True)
        L["e1b"] = Conv3(c1, c1, True)
        L["e2a"] = Conv3(c1, c2, True)
        L["e2b"] = Conv3(c2, c2, True)
        L["e3a"] = Conv3(c2, c3, True)
        L["e3b"] = Conv3(c3, c3, True)
        L["b1"] = Conv3(c3, c4, True)
        L["b2"] = Conv3(c4, c4, True)
        L["b3"] = Conv3(c4, c4, True)
        L["u3"] = Conv3(c4, c3, True)
        L["d3"] = Conv3(c3, c3, True)
        L["u2"] = Conv3(c3, c2, True)
        L["d2"] = Conv3(c2, c2, True)
        L["u1"] = Conv3(c2, c1, True)
        L["d1"] = Conv3(c1, c1, True)
        L["head"] = Conv3(c1, 8, False, out_f32=True)
        self.layers = L
        sizes = []
        for conv in L.values():
            conv.wi = len(sizes)
            conv.bi = len(sizes) + 1
            sizes += conv.sizes()
            conv.cout_live = conv.cout
        L["head"].cout_live = 1
        self.params = _Params(sizes, self.device)
        self.kappa = 1.0
        self.work = _Work(self.device)
        self._init(seed)

    # ---------------------------------------------------------------- set-up
    @property
    def n_params(self) -> int:
        """Live parameters (padded input channels of the first layer and the
        head's padded output channels excluded)."""
        n = 0
        for name, conv in self.layers.items():
            cin = 1 if name == "e1a" else conv.cin
            n += conv.cout_live * 9 * cin + conv.cout_live
        return n

    def _init(self, seed):
        """He-normal weights on the live entries, zero biases."""
        import torch
        rng = np.random.default_rng(seed)
        w = np.zeros(self.params.n, dtype=np.float32)
        for name, conv in self.layers.items():
            cin = 1 if name == "e1a" else conv.cin
            std = math.sqrt(2.0 / (9 * cin)) if conv.relu else math.sqrt(1.0 / (9 * cin))
            a = int(self.params.offsets[conv.wi])
            W = np.zeros((conv.cout, 9, conv.cin), dtype=np.float32)
            W[:conv.cout_live, :, :cin] = rng.standard_normal((conv.cout_live, 9, cin)) * std
            w[a:a + W.size] = W.reshape(-1)
        self.params.w.copy_(torch.as_tensor(w, device=self.device))
        self.params.refresh_bf16()

    def state_dict(self):
        return {"w": self.params.w.detach().cpu().numpy().copy(),
                "channels": np.asarray(self.channels, dtype=np.int64),
                "kappa": np.asarray([self.kappa], dtype=np.float64)}

    def load_state_dict(self, sd):
        import torch
        if tuple(int(c) for c in sd["channels"]) != self.channels:
            raise ValueError("channel layout mismatch")
        w = np.asarray(sd["w"], dtype=np.float32)
        if w.size != self.params.n:
            raise ValueError("parameter count mismatch")
        self.params.w.copy_(torch.as_tensor(w, device=self.device))
        self.kappa = float(np.asarray(sd["kappa"]).reshape(-1)[0])
        self.params.refresh_bf16()

    def save(self, path):
        from safetensors.numpy import save_file
        save_file({k: np.ascontiguousarray(v) for k, v in self.state_dict().items()}, str(path))

    @classmethod
    def load(cls, path, device=None):
        from safetensors.numpy import load_file
        sd = load_file(str(path))
        net = cls(tuple(int(c) for c in sd["channels"]), device=device)
        net.load_state_dict(sd)
        return net

    # ------------------------------------------------------------ primitives
    def _conv_fwd(self, conv: Conv3, x, B, H, W):
        import torch
        Wb = self.params.view(conv.wi, (conv.cout, 9 * conv.cin), bf16=True)
        bias = self.params.view(conv.bi, (conv.cout,))
        dt = torch.float32 if conv.out_f32 else torch.bfloat16
        y = torch.empty((B, H, W, conv.cout), dtype=dt, device=self.device)
        _call("bh_nn_conv3", _lib.ptr(x), B, H, W, conv.cin, _lib.ptr(Wb), conv.cout,
              _lib.ptr(bias), int(conv.relu), 1.0, _lib.ptr(y), int(not conv.out_f32), _st())
        return y

    def _conv_bwd(self, conv: Conv3, x, y, dy, B, H, W, need_dx=True):
        """dy: gradient of the layer output (bf16).  Accumulates nothing:
        writes the weight and bias gradients into the flat gradient buffer
        and returns dx (bf16) or None."""
        import torch
        P = B * H * W
        n = P * conv.cout
        if conv.relu:
            dz = self.work.get("dz", n, torch.bfloat16)
            _call("bh_nn_relu_bwd", _lib.ptr(dy), _lib.ptr(y), n, _lib.ptr(dz), _st())
        else:
            dz = dy.reshape(-1)
        K = 9 * conv.cin
        dW = self.params.view(conv.wi, (conv.cout, K), grad=True)
        ws_n = _lib.load().bh_nn_conv3_dw_workspace_size(B, H, W, conv.cin, conv.cout) // 4
        ws = self.work.get("split", max(ws_n, 1), torch.float32)
        _call("bh_nn_conv3_dw", _lib.ptr(x), _lib.ptr(dz), B, H, W, conv.cin, conv.cout,
              _lib.ptr(dW), _lib.ptr(ws), ws.numel() * 4, _st())
        db = self.params.view(conv.bi, (conv.cout,), grad=True)
        cs_n = _lib.load().bh_nn_colsum_workspace_size(P, conv.cout) // 4
        cs = self.work.get("colsum", max(cs_n, 1), torch.float32)
        _call("bh_nn_colsum", _lib.ptr(dz), P, conv.cout, _lib.ptr(db), _lib.ptr(cs),
              cs.numel() * 4, _st())
        if not need_dx:
            return None
        Wb = self.params.view(conv.wi, (conv.cout, K), bf16=True)
        Wf = self.work.get("wflip", conv.cin * 9 * conv.cout, torch.bfloat16)
        _call("bh_nn_conv3_wflip", _lib.ptr(Wb), conv.cout, conv.cin, _lib.ptr(Wf), _st())
        dx = torch.empty((B, H, W, conv.cin), dtype=torch.bfloat16, device=self.device)
        _call("bh_nn_conv3", _lib.ptr(dz), B, H, W, conv.cout, _lib.ptr(Wf), conv.cin, None, 0,
              1.0, _lib.ptr(dx), 1, _st())
        return dx

    def _pool(self, x, B, H, W, Cc):
        import torch
        y = torch.empty((B, H // 2, W // 2, Cc), dtype=torch.bfloat16, device=self.device)
        _call("bh_nn_pool2", _lib.ptr(x), B, H, W, Cc, _lib.ptr(y), 0, _st())
        return y

    def _pool_bwd(self, dy, B, H, W, Cc):
        import torch
        dx = torch.empty((B, H, W, Cc), dtype=torch.bfloat16, device=self.device)
        _call("bh_nn_pool2", _lib.ptr(dy), B, H, W, Cc, _lib.ptr(dx), 1, _st())
        return dx

    def _up(self, x, B, H, W, Cc):
        import torch
        y = torch.empty((B, 2 * H, 2 * W, Cc), dtype=torch.bfloat16, device=self.device)
        _call("bh_nn_up2", _lib.ptr(x), B, H, W, Cc, _lib.ptr(y), 0, _st())
        return y

    def _up_bwd(self, dy, B, H, W, Cc):
        import torch
        dx = torch.empty((B, H, W, Cc), dtype=torch.bfloat16, device=self.device)
        _call("bh_nn_up2", _lib.ptr(dy), B, H, W, Cc, _lib.ptr(dx), 1, _st())
        return dx

    def _add(self, a, b):
        import torch
        z = torch.empty_like(a)
        _call("bh_nn_add", _lib.ptr(a), _lib.ptr(b), a.numel(), _lib.ptr(z), _st())
        return z

    # ------------------------------------------------------------- the model
    def input_tensor(self, images):
        """(B, H, W) float images -> the (B, H, W, 8) bf16 network input with
        the image in channel 0 (already normalised by the caller)."""
        import torch
        B, H, W = images.shape
        x = torch.zeros((B, H, W, 8), dtype=torch.bfloat16, device=self.device)
        x[..., 0] = images.to(device=self.device, dtype=torch.float32).to(torch.bfloat16)
        return x

    def forward(self, x):
        """x: (B, H, W, 8) bf16, H and W multiples of 8.  Returns the FP32
        output (B, H, W) and keeps the activations for ``backward``."""
        B, H, W, _ = x.shape
        if H % 8 or W % 8:
            raise ValueError("image sides must be multiples of 8")
        L = self.layers
        c1, c2, c3, c4 = self.channels
        a = {"x": x}
        a["e1a"] = self._conv_fwd(L["e1a"], x, B, H, W)
        a["s1"] = self._conv_fwd(L["e1b"], a["e1a"], B, H, W)
        a["p1"] = self._pool(a["s1"], B, H, W, c1)
        H2, W2 = H // 2, W // 2
        a["e2a"] = self._conv_fwd(L["e2a"], a["p1"], B, H2, W2)
        a["s2"] = self._conv_fwd(L["e2b"], a["e2a"], B, H2, W2)
        a["p2"] = self._pool(a["s2"], B, H2, W2, c2)
        H3, W3 = H2 // 2, W2 // 2
        a["e3a"] = self._conv_fwd(L["e3a"], a["p2"], B, H3, W3)
        a["s3"] = self._conv_fwd(L["e3b"], a["e3a"], B, H3, W3)
        a["p3"] = self._pool(a["s3"], B, H3, W3, c3)
        H4, W4 = H3 // 2, W3 // 2
        a["b1"] = self._conv_fwd(L["b1"], a["p3"], B, H4, W4)
        a["b2"] = self._conv_fwd(L["b2"], a["b1"], B, H4, W4)
        a["b3"] = self._conv_fwd(L["b3"], a["b2"], B, H4, W4)
        a["up3"] = self._up(a["b3"], B, H4, W4, c4)
        a["u3"] = self._conv_fwd(L["u3"], a["up3"], B, H3, W3)
        a["a3"] = self._add(a["u3"], a["s3"])
        a["d3"] = self._conv_fwd(L["d3"], a["a3"], B, H3, W3)
        a["up2"] = self._up(a["d3"], B, H3, W3, c3)
        a["u2"] = self._conv_fwd(L["u2"], a["up2"], B, H2, W2)
        a["a2"] = self._add(a["u2"], a["s2"])
        a["d2"] = self._conv_fwd(L["d2"], a["a2"], B, H2, W2)
        a["up1"] = self._up(a["d2"], B, H2, W2, c2)
        a["u1"] = self._conv_fwd(L["u1"], a["up1"], B, H, W)
        a["a1"] = self._add(a["u1"], a["s1"])
        a["d1"] = self._conv_fwd(L["d1"], a["a1"], B, H, W)
        a["out"] = self._conv_fwd(L["head"], a["d1"], B, H, W)      # FP32 (B, H, W, 8)
        self._acts = (a, B, H, W)
        return a["out"][..., 0]

    def backward(self, grad_out):
        """grad_out: (B, H, W, 8) bf16 gradient of the FP32 head output (only
        channel 0 non-zero).  Fills the flat gradient buffer."""
        a, B, H, W = self._acts
        L = self.layers
        c1, c2, c3, c4 = self.channels
        H2, W2, H3, W3, H4, W4 = H // 2, W // 2, H // 4, W // 4, H // 8, W // 8
        g = self._conv_bwd(L["head"], a["d1"], a["out"], grad_out, B, H, W)
        g_a1 = self._conv_bwd(L["d1"], a["a1"], a["d1"], g, B, H, W)
        g = self._conv_bwd(L["u1"], a["up1"], a["u1"], g_a1, B, H, W)
        g = self._up_bwd(g, B, H2, W2, c2)
        g_a2 = self._conv_bwd(L["d2"], a["a2"], a["d2"], g, B, H2, W2)
        g = self._conv_bwd(L["u2"], a["up2"], a["u2"], g_a2, B, H2, W2)
        g = self._up_bwd(g, B, H3, W3, c3)
        g_a3 = self._conv_bwd(L["d3"], a["a3"], a["d3"], g, B, H3, W3)
        g = self._conv_bwd(L["u3"], a["up3"], a["u3"], g_a3, B, H3, W3)
        g = self._up_bwd(g, B, H4, W4, c4)
        g = self._conv_bwd(L["b3"], a["b2"], a["b3"], g, B, H4, W4)
        g = self._conv_bwd(L["b2"], a["b1"], a["b2"], g, B, H4, W4)
        g = self._conv_bwd(L["b1"], a["p3"], a["b1"], g, B, H4, W4)
        g = self._pool_bwd(g, B, H3, W3, c3)
        g = self._add(g, g_a3)                        # s3 feeds the pool and the skip
        g = self._conv_bwd(L["e3b"], a["e3a"], a["s3"], g, B, H3, W3)
        g = self._conv_bwd(L["e3a"], a["p2"], a["e3a"], g, B, H3, W3)
        g = self._pool_bwd(g, B, H2, W2, c2)
        g = self._add(g, g_a2)
        g = self._conv_bwd(L["e2b"], a["e2a"], a["s2"], g, B, H2, W2)
        g = self._conv_bwd(L["e2a"], a["p1"], a["e2a"], g, B, H2, W2)
        g = self._pool_bwd(g, B, H, W, c1)
        g = self._add(g, g_a1)
        g = self._conv_bwd(L["e1b"], a["e1a"], a["s1"], g, B, H, W)
        self._conv_bwd(L["e1a"], a["x"], a["e1a"], g, B, H, W, need_dx=False)

    def loss_and_grad(self, target):
        """Squared-error loss of the last forward output (channel 0) against
        target (B, H, W) FP32; returns (loss, grad_out)."""
        import torch
        a, B, H, W = self._acts
        P = B * H * W
        t = target.to(device=self.device, dtype=torch.float32).contiguous()
        grad = torch.empty((B, H, W, 8), dtype=torch.bfloat16, device=self.device)
        nb = (P + 255) // 256
        part = self.work.get("mse", nb, torch.float32)
        _call("bh_nn_mse", _lib.ptr(a["out"]), 8, _lib.ptr(t), P, _lib.ptr(grad), 8,
              _lib.ptr(part), nb * 4, _st())
        return part.sum() / P, grad

    def step(self, images, targets, lr):
        """One training step on a batch; returns the loss (device scalar)."""
        self.forward(self.input_tensor(images))
        loss, grad = self.loss_and_grad(targets)
        self.backward(grad)
        self.params.adam(lr)
        return loss

    # ------------------------------------------------------------ inference
    def predict(self, image):
        """p0 estimate from one adjoint image (H, W) (any scale)."""
        import torch
        img = image.to(device=self.device, dtype=torch.float64)
        rms = float(img.pow(2).mean().sqrt().item()) or 1.0
        out = self.forward(self.input_tensor((img / rms).reshape((1,) + tuple(img.shape))))
        self._acts = None
        return out[0].to(torch.float64) * (rms / self.kappa)


# ---------------------------------------------------------------- training
@dataclass
class PairSet:
    """(adjoint image, p0) pairs on the device, FP32."""
    images: "object"
    targets: "object"
    seeds: List[int] = field(default_factory=list)
    seconds: float = 0.0


def recon_problem(device, N=(256, 256), dx=(1e-4, 1e-4)):
    """The reconstruction workload: 256^2, dx 1e-4, c 1500, cfl 0.3, 3 levels
    (4,8,16) rho 2, four-edge sensors, top_n 32768, FP32-split."""
    import beamax_amd as bx
    from .gb import gb_solvers
    from .solvers.faces import FourEdgeAdjoint, four_edge_mask
    dom = bx.Domain(N=N, dx=dx, c=1500.0, cfl=0.3, periodic=(False, False))
    wpt = bx.MSWPT(bx.DyadicDecomposition(3, N, (4, 8, 16), (1, 1)), 2, "rectangular_mirror",
                   device=device)
    sensors = bx.Sensor(dom, binary_mask=four_edge_mask(N))
    ts = dom.generate_time_domain()
    s = bx.MSGBSolver(32768, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                      precision="fp32_split", cull_log2=40.0)
    fa = FourEdgeAdjoint(s, dom, ts, 512, device=device)
    return dom, wpt, sensors, ts, s, fa


def measured_data(problem, p0, seed, noise=0.01):
    """MSGB forward data of p0 plus Gaussian noise of ``noise`` x the data RMS."""
    import torch
    dom, wpt, sensors, ts, s, fa = problem
    d = s.forward(p0, dom, sensors, ts, wpt).to(torch.float64)
    gen = torch.Generator(device="cpu").manual_seed(int(seed))
    return d + noise * d.pow(2).mean().sqrt() * torch.randn(d.shape, generator=gen,
                                                            dtype=torch.float64).to(d.device)


def make_pairs(problem, seeds, device, log=None):
    """Adjoint image of noisy MSGB data and the phantom, for each seed
    (``workloads.phantom_2d``)."""
    import torch
    from .workloads import phantom_2d
    dom, wpt, sensors, ts, s, fa = problem
    N = dom.N
    imgs = torch.empty((len(seeds),) + tuple(N), dtype=torch.float32, device=device)
    tgts = torch.empty_like(imgs)
    t0 = time.perf_counter()
    for i, sd in enumerate(seeds):
        p0 = torch.as_tensor(phantom_2d(N, dom.dx, seed=int(sd)), device=device)
        d = measured_data(problem, p0, sd)
        imgs[i] = fa(d).reshape(N).to(torch.float32)
        tgts[i] = p0.to(torch.float32)
        if log is not None and (i + 1) % 256 == 0:
            log(f"pairs {i + 1}/{len(seeds)} {time.perf_counter() - t0:.1f}s")
    torch.cuda.synchronize()
    return PairSet(imgs, tgts, list(seeds), time.perf_counter() - t0)


def normalise(pairs: PairSet):
    """Scale each input to unit RMS and its target by the same factor; return
    kappa = 1 / median(RMS(target) / RMS(input)) so targets are O(1)."""
    import torch
    r_in = pairs.images.pow(2).mean(dim=(1, 2)).sqrt().clamp_min(1e-30)
    ratio = pairs.targets.pow(2).mean(dim=(1, 2)).sqrt() / r_in
    kappa = 1.0 / float(ratio.median().item())
    x = pairs.images / r_in[:, None, None]
    y = pairs.targets / r_in[:, None, None] * kappa
    return x, y, kappa


def train(net: UNet, pairs: PairSet, *, epochs=8, batch=16, lr=1e-3, seed=0, log=None):
    """Adam on the squared error, cosine learning-rate decay; returns the
    per-epoch mean losses."""
    import torch
    x, y, kappa = normalise(pairs)
    net.kappa = kappa
    n = x.shape[0]
    rng = np.random.default_rng(seed)
    steps = epochs * (n // batch)
    k = 0
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = torch.as_tensor(rng.permutation(n), device=x.device)
        tot = torch.zeros((), dtype=torch.float32, device=x.device)
        for b in range(n // batch):
            idx = perm[b * batch:(b + 1) * batch]
            lr_k = lr * 0.5 * (1.0 + math.cos(math.pi * k / max(steps, 1)))
            tot += net.step(x[idx], y[idx], lr_k)
            k += 1
        hist.append(float(tot.item()) / max(n // batch, 1))
        if log is not None:
            log(f"epoch {ep + 1}/{epochs} loss {hist[-1]:.5f} {time.perf_counter() - t0:.1f}s")
    return hist


def warm_start(net: UNet, op, data, image):
    """x0 = alpha * net(image) with alpha the least-squares data fit
    <F x, d> / <F x, F x> (one operator application)."""
    import torch
    shape = tuple(op.shape[1])
    xh = net.predict(image.reshape(shape)).reshape(shape)
    w = op.apply(xh).to(torch.float64)
    d = data.reshape(w.shape).to(torch.float64)
    ww = float((w * w).sum().item())
    alpha = float((w * d).sum().item()) / ww if ww > 0 else 1.0
    return alpha * xh


def evaluate(net: Optional[UNet], problem, seeds, device, tol=1e-3, log=None, graph=True):
    """CG on the normal equations from zero and from the learned start, each
    also with the Fourier preconditioner (``reconstruct.FourierPreconditioner``,
    |k|^-0.5 with floor 0.03), for each held-out seed: iterations, seconds
    (CG alone, and with the network's inference and the least-squares
    scaling), error against p0.  Every run replays a captured CG iteration
    (``graph``)."""
    import torch
    from .reconstruct import FourierPreconditioner, cgnr
    from .solvers.linear import FrozenForward
    from .workloads import phantom_2d
    dom, wpt, sensors, ts, s, fa = problem
    N = dom.N
    pc = FourierPreconditioner(N, device=device)
    rows = []
    for sd in seeds:
        p0 = torch.as_tensor(phantom_2d(N, dom.dx, seed=int(sd)), device=device)
        d = measured_data(problem, p0, sd)
        img = fa(d).reshape(N)
        F = FrozenForward.from_field(s, img, dom, sensors, ts, wpt)
        cgnr(F, d, tol=tol, maxiter=3, graph=graph)      # warm-up (workspaces)
        x, info = cgnr(F, d, tol=tol, maxiter=1000, graph=graph)
        row = {"seed": int(sd), "plain_iters": info["iterations"],
               "plain_s": info["seconds"],
               "plain_err": float(((x - p0).norm() / p0.norm()).item())}
        xp, infop = cgnr(F, d, tol=tol, maxiter=1000, precond=pc, graph=graph)
        row.update({"pc_iters": infop["iterations"], "pc_s": infop["seconds"],
                    "pc_err": float(((xp - p0).norm() / p0.norm()).item())})
        if net is not None:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x0 = warm_start(net, F, d, img)
            torch.cuda.synchronize()
            t_init = time.perf_counter() - t0
            x1, info1 = cgnr(F, d, x0=x0, tol=tol, maxiter=1000, graph=graph)
            row.update({"learned_iters": info1["iterations"], "learned_cg_s": info1["seconds"],
                        "learned_s": info1["seconds"] + t_init, "init_s": t_init,
                        "start_residual": info1["residuals"][0],
                        "start_err": float(((x0 - p0).norm() / p0.norm()).item()),
                        "learned_err": float(((x1 - p0).norm() / p0.norm()).item())})
            x2, info2 = cgnr(F, d, x0=x0, tol=tol, maxiter=1000, precond=pc, graph=graph)
            row.update({"learned_pc_iters": info2["iterations"],
                        "learned_pc_s": info2["seconds"] + t_init,
                        "learned_pc_err": float(((x2 - p0).norm() / p0.norm()).item())})
        rows.append(row)
        if log is not None:
            log(" ".join(f"{k}={v:.4g}" if isinstance(v, float) else f"{k}={v}"
                         for k, v in row.items()))
    return rows


def conv_layer_timing(H: int, cin: int, cout: int, batch: int = 16, reps: int = 3,
                      device=None):
    """Forward and weight gradient of one 3x3 layer (cin -> cout channels,
    batch x H x H pixels, random bf16 data), each run ``reps`` times on the
    current stream after one untimed call; returns the mean microseconds per
    call of each (HIP events).  A profiling target for the two convolution
    kernels on their own."""
    import torch
    dev = torch.device("cuda") if device is None else torch.device(device)
    g = torch.Generator(device=dev).manual_seed(0)
    P = batch * H * H
    x = torch.randn(P * cin, device=dev, generator=g).to(torch.bfloat16)
    w = torch.randn(cout * 9 * cin, device=dev, generator=g).to(torch.bfloat16)
    dz = torch.randn(P * cout, device=dev, generator=g).to(torch.bfloat16)
    y = torch.empty(P * cout, device=dev, dtype=torch.bfloat16)
    dW = torch.empty(cout * 9 * cin, device=dev, dtype=torch.float32)
    nb = _lib.load().bh_nn_conv3_dw_workspace_size(batch, H, H, cin, cout)
    ws = torch.empty(max(nb // 4, 1), device=dev, dtype=torch.float32)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd = dw = 0.0
    for it in range(reps + 1):   # the first, untimed, loads the kernels
        ev[0].record()
        _call("bh_nn_conv3", _lib.ptr(x), batch, H, H, cin, _lib.ptr(w), cout, None, 1, 1.0,
              _lib.ptr(y), 1, _st())
        ev[1].record()
        _call("bh_nn_conv3_dw", _lib.ptr(x), _lib.ptr(dz), batch, H, H, cin, cout, _lib.ptr(dW),
              _lib.ptr(ws), nb, _st())
        ev[2].record()
        torch.cuda.synchronize(dev)
        if it > 0:
            fwd += ev[0].elapsed_time(ev[1]) * 1e3 / reps
            dw += ev[1].elapsed_time(ev[2]) * 1e3 / reps
    return {"forward_us": fwd, "weight_grad_us": dw}


def _main(argv=None):
    """Train the network on generated pairs, save it, and compare CG with and
    without it on held-out seeds (one JSON line on stdout)."""
    import argparse
    import json
    import torch
    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=32, help="held-out seeds 0 .. n-1")
    ap.add_argument("--out", default=None, help="weights file (safetensors)")
    ap.add_argument("--weights", default=None, help="evaluate these weights, no training")
    ap.add_argument("--conv", type=int, nargs=3, metavar=("H", "CIN", "COUT"), default=None,
                    help="only time the forward and weight gradient of one 3x3 layer")
    a = ap.parse_args(argv)
    dev = torch.device("cuda")
    if a.conv:
        H, cin, cout = a.conv
        print(json.dumps({"layer": [H, cin, cout], "batch": a.batch,
                          **conv_layer_timing(H, cin, cout, a.batch, reps=10, device=dev)}),
              flush=True)
        return
    log = lambda m: print(m, flush=True)
    problem = recon_problem(dev)
    result = {}
    if a.weights:
        net = UNet.load(a.weights, device=dev)
    else:
        seeds = list(range(1000, 1000 + a.pairs))
        pairs = make_pairs(problem, seeds, dev, log=log)
        log(f"pairs done: {len(seeds)} in {pairs.seconds:.1f}s")
        net = UNet(device=dev)
        t0 = time.perf_counter()
        hist = train(net, pairs, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
        torch.cuda.synchronize()
        result.update({"train_pairs": len(seeds), "pair_s": round(pairs.seconds, 2),
                       "epochs": a.epochs, "batch": a.batch,
                       "train_s": round(time.perf_counter() - t0, 2),
                       "loss_per_epoch": [round(h, 6) for h in hist]})
        del pairs
        if a.out:
            net.save(a.out)
            log(f"saved {a.out}")
    result["params"] = net.n_params
    if a.eval <= 0:
        print(json.dumps(result), flush=True)
        return
    rows = evaluate(net, problem, list(range(a.eval)), dev, log=log)
    mean = lambda k: float(np.mean([r[k] for r in rows]))
    ratios = np.array([r["learned_s"] / r["plain_s"] for r in rows])
    result["time_ratio_per_seed"] = {
        "median": float(np.median(ratios)), "min": float(ratios.min()),
        "max": float(ratios.max()), "p10": float(np.percentile(ratios, 10)),
        "p90": float(np.percentile(ratios, 90)),
        "iterations_plain_hist": np.bincount([r["plain_iters"] for r in rows]).tolist(),
        "iterations_learned_hist": np.bincount([r["learned_iters"] for r in rows]).tolist()}
    ratios_pc = np.array([r["learned_pc_s"] / r["plain_s"] for r in rows])
    result["learned_pc_time_ratio_per_seed"] = {
        "median": float(np.median(ratios_pc)), "min": float(ratios_pc.min()),
        "max": float(ratios_pc.max()), "p10": float(np.percentile(ratios_pc, 10)),
        "p90": float(np.percentile(ratios_pc, 90)),
        "iterations_hist": np.bincount([r["learned_pc_iters"] for r in rows]).tolist()}
    result.update({"pc_iters": mean("pc_iters"), "pc_s": mean("pc_s"), "pc_err": mean("pc_err"),
                   "learned_pc_iters": mean("learned_pc_iters"),
                   "learned_pc_s": mean("learned_pc_s"), "learned_pc_err": mean("learned_pc_err"),
                   "pc_time_ratio": mean("pc_s") / mean("plain_s"),
                   "learned_pc_time_ratio": mean("learned_pc_s") / mean("plain_s")})
    result.update({"eval_seeds": len(rows), "plain_iters": mean("plain_iters"),
                   "plain_s": mean("plain_s"), "learned_iters": mean("learned_iters"),
                   "learned_s": mean("learned_s"), "learned_cg_s": mean("learned_cg_s"),
                   "init_s": mean("init_s"), "start_residual": mean("start_residual"),
                   "start_err": mean("start_err"), "plain_err": mean("plain_err"),
                   "learned_err": mean("learned_err"),
                   "time_ratio": mean("learned_s") / mean("plain_s")})
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    _main()

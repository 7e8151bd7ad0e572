"""Full-state surrogates: ResNet, U-Net and FNO mapping one normalised state (B, 4, n, n) to the next.

The architectures of the reference's surrogate/models.py (Equinox), rebuilt on torch for the GPU:

  ResNet  (models.py:48-79)    1x1 lift -> `depth` pre-activation residual blocks of two k x k periodic
                               convolutions -> 1x1 projection; predicts the one-frame change,
                               next = x + delta_scale * net(x); zero-initialised projection.
  UNet    (models.py:80-152)   ConvBlocks down `levels` halvings (2x2 average pooling) and back up
                               (nearest-neighbour upsampling, skip concatenation); channels double per level.
  FNO     (models.py:155-248)  spectral convolutions keeping |kx| < modes, 0 <= ky < modes (two complex weight
                               blocks for kx >= 0 and kx < 0) plus a 1x1 path, GELU; lift / project 1x1.
  Positive (models.py:259-300) wrapper working in (log rho, u, v, log p): density and pressure of the output
                               are positive by construction.

All convolutions use periodic padding, so the ResNet and FNO are exactly translation-equivariant on the
torus (the U-Net for shifts that are multiples of 2**levels).  ResNet(backend="mfma") and UNet(backend="mfma")
run their 3x3 convolutions on the library's BF16-MFMA periodic convolution kernels (npde_conv3x3 and its
weight gradient; forward and backward) in NHWC, with the GELU that follows the first convolution of each
block in the kernel's epilogue; the 1x1 convolutions are matrix products and the U-Net's pooling,
upsampling and skip concatenation tensor operations on the NHWC activations.  FNO(backend="dft") computes
its spectral layers as truncated discrete Fourier transforms in matrix products (only the kept modes are
formed) instead of rfft2 / irfft2.  Each backend uses the torch backend's parameters, so a model can switch.  GELU is the tanh approximation (JAX's
default, which the reference uses).  Parameter shapes follow the reference, so parameter counts agree:
ResNet(64, 8) 591,428; UNet(16, 3) 516,516; FNO(32, 4, 16 modes) 4,203,428.
"""
from __future__ import annotations

import math

import torch
import torch.nn.functional as F

GAMMA = 1.4
TINY = 1e-6


def gelu(x):
    return F.gelu(x, approximate="tanh")


class PeriodicConv(torch.nn.Module):
    """k x k convolution with wrap-around padding (models.py:16-29).  Weights are initialised like
    Equinox's Conv2d: uniform in +-1/sqrt(fan_in)."""

    def __init__(self, c_in: int, c_out: int, k: int, generator: torch.Generator | None = None):
        super().__init__()
        bound = 1.0 / math.sqrt(c_in * k * k)
        self.weight = torch.nn.Parameter((torch.rand(c_out, c_in, k, k, generator=generator) * 2 - 1) * bound)
        self.bias = torch.nn.Parameter((torch.rand(c_out, generator=generator) * 2 - 1) * bound)
        self.pad = k // 2

    def zero_(self):
        with torch.no_grad():
            self.weight.zero_()
            self.bias.zero_()
        return self

    def forward(self, x):
        if self.pad:
            p = self.pad
            x = F.pad(x, (p, p, p, p), mode="circular")
        return F.conv2d(x, self.weight, self.bias)


def _scale(delta_scale, channels):
    return torch.tensor((1.0,) * channels if delta_scale is None else tuple(float(d) for d in delta_scale))


class ResBlock(torch.nn.Module):
    """x + conv(gelu(conv(gelu(x))))."""

    def __init__(self, width, k, generator=None):
        super().__init__()
        self.conv1 = PeriodicConv(width, width, k, generator)
        self.conv2 = PeriodicConv(width, width, k, generator)

    def forward(self, x):
        return x + self.conv2(gelu(self.conv1(gelu(x))))


class Conv3x3MFMA(torch.autograd.Function):
    """One 3x3 periodic convolution on the library's BF16-MFMA kernels, NHWC float32 (B, H, W, C) with C a
    multiple of 16 (at most 64), optionally followed by tanh-GELU in the kernel's epilogue.  Backward: the
    pre-activation gradient (GELU' of the saved pre-activation), the weight and bias gradients
    (npde_conv3x3_wgrad, fixed-order band reduction) and the data gradient (the transposed, tap-reversed
    convolution)."""

    @staticmethod
    def forward(ctx, x, weight, bias, apply_gelu):
        from .learned import conv as K
        cout = weight.shape[0]
        y, pre = K.conv3x3(x.contiguous(), K.pack_weights(weight.detach()), bias.detach().contiguous(), cout, cout,
                           K.EPI_GELU if apply_gelu else K.EPI_LINEAR, want_pre=apply_gelu)
        ctx.save_for_backward(x, weight, pre)
        ctx.apply_gelu = apply_gelu
        return y

    @staticmethod
    def backward(ctx, gy):
        from .learned import conv as K
        x, weight, pre = ctx.saved_tensors
        g = gy.contiguous()
        if ctx.apply_gelu:
            with torch.enable_grad():
                z = pre.detach().requires_grad_(True)
                g, = torch.autograd.grad(gelu(z), z, g)
        dw, db = K.conv3x3_wgrad(x.contiguous(), g.contiguous())
        cin = x.shape[-1]
        gx, _ = K.conv3x3(g.contiguous(), K.pack_weights(weight.detach(), transpose_flip=True), None, cin, cin,
                          K.EPI_LINEAR)
        return gx, dw.contiguous(), db, None


def _conv1x1_nhwc(h, conv: "PeriodicConv"):
    """A 1x1 PeriodicConv applied to NHWC activations: a matrix product over the channel axis."""
    return torch.matmul(h, conv.weight[:, :, 0, 0].t()) + conv.bias


class ResNet(torch.nn.Module):
    def __init__(self, width=64, depth=8, kernel_size=3, delta_scale=None, channels=4, seed=0, backend="torch"):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.lift = PeriodicConv(channels, width, 1, g)
        self.blocks = torch.nn.ModuleList([ResBlock(width, kernel_size, g) for _ in range(depth)])
        self.project = PeriodicConv(width, channels, 1, g).zero_()
        self.register_buffer("delta_scale", _scale(delta_scale, channels))
        self.kernel_size = kernel_size
        self.backend = backend

    def forward(self, x):
        if self.backend == "mfma":
            return self._forward_mfma(x)
        h = self.lift(x)
        for b in self.blocks:
            h = b(h)
        return x + self.delta_scale.view(1, -1, 1, 1).to(x.dtype) * self.project(gelu(h))

    def _forward_mfma(self, x):
        """The same network in NHWC with the 3x3 convolutions on the MFMA kernels (float32 activations,
        BF16 products, FP32 accumulation)."""
        width = self.lift.weight.shape[0]
        H, W = x.shape[-2:]
        if self.kernel_size != 3 or width % 16 or width > 128 or W % 8 or H % 8 or x.dtype != torch.float32:
            raise ValueError("the MFMA backend needs k = 3, width a multiple of 16 up to 128, H and W multiples of "
                             "8 and float32 input")
        h = _conv1x1_nhwc(x.permute(0, 2, 3, 1), self.lift)
        for b in self.blocks:
            t = Conv3x3MFMA.apply(gelu(h), b.conv1.weight, b.conv1.bias, True)
            h = h + Conv3x3MFMA.apply(t, b.conv2.weight, b.conv2.bias, False)
        out = _conv1x1_nhwc(gelu(h), self.project).permute(0, 3, 1, 2)
        return x + self.delta_scale.view(1, -1, 1, 1).to(x.dtype) * out


class ConvBlock(torch.nn.Module):
    """skip(x) + conv(gelu(conv(gelu(x)))), with a 1x1 skip when the channel count changes."""

    def __init__(self, c_in, c_out, generator=None):
        super().__init__()
        self.conv1 = PeriodicConv(c_in, c_out, 3, generator)
        self.conv2 = PeriodicConv(c_out, c_out, 3, generator)
        self.skip = None if c_in == c_out else PeriodicConv(c_in, c_out, 1, generator)

    def forward(self, x):
        h = self.conv2(gelu(self.conv1(gelu(x))))
        return (x if self.skip is None else self.skip(x)) + h


def downsample(x):
    """2x2 average pooling (the dataset's own coarsening)."""
    B, c, n, m = x.shape
    return x.reshape(B, c, n // 2, 2, m // 2, 2).mean(dim=(3, 5))


def upsample(x):
    """Nearest-neighbour upsampling by 2."""
    return x.repeat_interleave(2, dim=-2).repeat_interleave(2, dim=-1)


class UNet(torch.nn.Module):
    def __init__(self, width=16, levels=3, delta_scale=None, channels=4, seed=0, backend="torch"):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        ch = [width * 2 ** l for l in range(levels + 1)]
        self.lift = PeriodicConv(channels, ch[0], 1, g)
        self.down = torch.nn.ModuleList([ConvBlock(ch[max(l - 1, 0)], ch[l], g) for l in range(levels)])
        self.bottleneck = ConvBlock(ch[levels - 1], ch[levels], g)
        self.up = torch.nn.ModuleList([ConvBlock(ch[l + 1] + ch[l], ch[l], g) for l in range(levels)])
        self.project = PeriodicConv(ch[0], channels, 1, g).zero_()
        self.register_buffer("delta_scale", _scale(delta_scale, channels))
        self.levels = levels
        self.backend = backend

    def _forward_mfma(self, x):
        """The same network in NHWC with every 3x3 convolution on the MFMA kernels (float32 activations, BF16
        products, FP32 accumulation); pooling, upsampling, concatenation and the 1x1 convolutions are tensor
        operations on the NHWC activations."""
        H, W = x.shape[-2:]
        widths = [self.lift.weight.shape[0] * 2 ** l for l in range(self.levels + 1)]
        if (any(c % 16 for c in widths) or widths[-1] > 128 or H % (8 << self.levels) or W % (8 << self.levels)
                or x.dtype != torch.float32):
            raise ValueError("the MFMA backend needs widths that are multiples of 16 up to 128 at the bottleneck, "
                             "H and W multiples of 8 * 2**levels and float32 input")

        def block(b, h):
            t = Conv3x3MFMA.apply(gelu(h), b.conv1.weight, b.conv1.bias, True)
            out = Conv3x3MFMA.apply(t, b.conv2.weight, b.conv2.bias, False)
            return (h if b.skip is None else _conv1x1_nhwc(h, b.skip)) + out

        def down(h):  # 2x2 average pooling
            B, n, m, c = h.shape
            return h.reshape(B, n // 2, 2, m // 2, 2, c).mean(dim=(2, 4))

        def up(h):  # nearest-neighbour upsampling by 2
            return h.repeat_interleave(2, dim=1).repeat_interleave(2, dim=2)

        h = _conv1x1_nhwc(x.permute(0, 2, 3, 1), self.lift)
        skips = []
        for b in self.down:
            h = block(b, h)
            skips.append(h)
            h = down(h)
        h = block(self.bottleneck, h)
        for b, skip in zip(reversed(self.up), reversed(skips)):
            h = block(b, torch.cat([up(h), skip], dim=-1))
        out = _conv1x1_nhwc(gelu(h), self.project).permute(0, 3, 1, 2)
        return x + self.delta_scale.view(1, -1, 1, 1).to(x.dtype) * out

    def forward(self, x):
        if self.backend == "mfma":
            return self._forward_mfma(x)
        h = self.lift(x)
        skips = []
        for block in self.down:
            h = block(h)
            skips.append(h)
            h = downsample(h)
        h = self.bottleneck(h)
        for block, skip in zip(reversed(self.up), reversed(skips)):
            h = block(torch.cat([upsample(h), skip], dim=1))
        return x + self.delta_scale.view(1, -1, 1, 1).to(x.dtype) * self.project(gelu(h))


class SpectralConv2d(torch.nn.Module):
    """Convolution as a multiplication of the lowest Fourier modes (models.py:155-190): rfft2, keep
    kx in [0, modes) and [-modes, 0), ky in [0, modes), mix channels with complex weights, irfft2."""

    def __init__(self, c_in, c_out, modes, generator=None):
        super().__init__()
        scale = 1.0 / (c_in * c_out)
        shape = (2, c_in, c_out, modes, modes)
        self.w_re = torch.nn.Parameter(scale * torch.rand(shape, generator=generator))
        self.w_im = torch.nn.Parameter(scale * torch.rand(shape, generator=generator))
        self.modes = modes
        self.backend = "torch"
        self._bases = {}

    def _basis(self, n, device, dtype):
        """cos / sin tables of the kept modes: columns ky in [0, m) for the real transform along the last axis,
        rows kx in [0, m) and [n - m, n) for the complex one along the other, in float64, cast once."""
        key = (n, device, dtype)
        if key not in self._bases:
            m = self.modes
            r = torch.arange(n, dtype=torch.float64)
            ky = torch.arange(m, dtype=torch.float64)
            kx = torch.cat([torch.arange(m), torch.arange(n - m, n)]).to(torch.float64)
            ay = 2 * math.pi * torch.outer(r, ky) / n   # (n, m)
            ax = 2 * math.pi * torch.outer(kx, r) / n   # (2m, n)
            cy = torch.full((m,), 2.0, dtype=torch.float64)
            cy[0] = 1.0                                 # one-sided spectrum: ky > 0 stands for +-ky
            t = {"cy": torch.cos(ay), "sy": torch.sin(ay), "cx": torch.cos(ax), "sx": torch.sin(ax),
                 "icy": (cy[:, None] * torch.cos(ay).t()) / (n * n), "isy": (cy[:, None] * torch.sin(ay).t()) / (n * n)}
            self._bases[key] = {k: v.to(device=device, dtype=dtype).contiguous() for k, v in t.items()}
        return self._bases[key]

    def _forward_dft(self, x):
        """The layer as a truncated discrete Fourier transform in matrix products (rocBLAS/hipBLASLt GEMMs):
        only the 2m x m kept modes are ever formed.  Forward along the last axis (real input, ky < m), then
        along the other (kx in [0, m) and [n - m, n)); the per-mode channel mixing as one batched product over
        the modes; back along the first axis, then the real inverse along the last, where a mode ky > 0 counts
        twice and the imaginary part of ky = 0 drops out (torch.fft.irfft2's one-sided convention).  Equal to
        the rfft2 path to rounding."""
        B, c_in, n, n2 = x.shape
        m = self.modes
        if n2 != n or 2 * m > n:
            raise ValueError("the DFT path needs square inputs and 2 * modes <= n")
        T = self._basis(n, x.device, x.dtype)
        ar, ai = x @ T["cy"], -(x @ T["sy"])                        # (B, ci, n, m): sum_y x e^{-i ky y}
        xr = torch.matmul(T["cx"], ar) + torch.matmul(T["sx"], ai)  # (B, ci, 2m, m): sum_x (.) e^{-i kx x}
        xi = torch.matmul(T["cx"], ai) - torch.matmul(T["sx"], ar)
        wr = torch.cat([self.w_re[0], self.w_re[1]], dim=2)        # (ci, co, 2m, m): kx in [0, m), then [n-m, n)
        wi = torch.cat([self.w_im[0], self.w_im[1]], dim=2)
        # per mode (kx, ky): Y[b, o] = sum_i X[b, i] W[i, o], complex; modes as the batch of one product
        Xr = xr.permute(2, 3, 0, 1).reshape(2 * m * m, B, c_in)
        Xi = xi.permute(2, 3, 0, 1).reshape(2 * m * m, B, c_in)
        Wr = wr.permute(2, 3, 0, 1).reshape(2 * m * m, c_in, -1).to(x.dtype)
        Wi = wi.permute(2, 3, 0, 1).reshape(2 * m * m, c_in, -1).to(x.dtype)
        yr = torch.bmm(Xr, Wr) - torch.bmm(Xi, Wi)
        yi = torch.bmm(Xr, Wi) + torch.bmm(Xi, Wr)
        c_out = yr.shape[-1]
        yr = yr.reshape(2 * m, m, B, c_out).permute(2, 3, 0, 1)   # (B, co, 2m, m)
        yi = yi.reshape(2 * m, m, B, c_out).permute(2, 3, 0, 1)
        # back along the first axis: sum_kx Y e^{+i kx x} (n, 2m) @ (2m, m); then the real inverse along the last
        zr = torch.matmul(T["cx"].t(), yr) - torch.matmul(T["sx"].t(), yi)  # (B, co, n, m)
        zi = torch.matmul(T["cx"].t(), yi) + torch.matmul(T["sx"].t(), yr)
        return zr @ T["icy"] - zi @ T["isy"]                        # (B, co, n, n)

    def forward(self, x):
        if self.backend == "dft":
            return self._forward_dft(x)
        B, c_in, n, _ = x.shape
        m = self.modes
        X = torch.fft.rfft2(x)
        W = torch.complex(self.w_re, self.w_im).to(X.dtype)
        Y = torch.zeros(B, W.shape[2], n, n // 2 + 1, dtype=X.dtype, device=x.device)
        Y[:, :, :m, :m] = torch.einsum("bixy,ioxy->boxy", X[:, :, :m, :m], W[0])
        Y[:, :, -m:, :m] = torch.einsum("bixy,ioxy->boxy", X[:, :, -m:, :m], W[1])
        return torch.fft.irfft2(Y, s=(n, n))


class FNOBlock(torch.nn.Module):
    def __init__(self, width, modes, generator=None):
        super().__init__()
        self.spectral = SpectralConv2d(width, width, modes, generator)
        self.pointwise = PeriodicConv(width, width, 1, generator)

    def forward(self, h):
        return gelu(self.spectral(h) + self.pointwise(h))


class FNO(torch.nn.Module):
    """backend "torch": spectral layers on torch.fft (rocFFT); "dft": on the truncated-DFT matrix products of
    SpectralConv2d._forward_dft.  The same parameters serve both."""

    def __init__(self, width=32, depth=4, modes=12, delta_scale=None, channels=4, hidden=128, seed=0,
                 backend="torch"):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.lift = PeriodicConv(channels, width, 1, g)
        self.blocks = torch.nn.ModuleList([FNOBlock(width, modes, g) for _ in range(depth)])
        if backend not in ("torch", "dft"):
            raise ValueError("FNO backend must be 'torch' or 'dft'")
        for b in self.blocks:
            b.spectral.backend = backend
        self.project1 = PeriodicConv(width, hidden, 1, g)
        self.project2 = PeriodicConv(hidden, channels, 1, g).zero_()
        self.register_buffer("delta_scale", _scale(delta_scale, channels))

    def forward(self, x):
        h = self.lift(x)
        for b in self.blocks:
            h = b(h)
        return x + self.delta_scale.view(1, -1, 1, 1).to(x.dtype) * self.project2(gelu(self.project1(h)))


def conserved_to_logprim(U):
    """(rho, rho u, rho v, E) -> (log rho, u, v, log p); U (B, 4, n, n) physical."""
    rho = torch.clamp(U[:, 0], min=TINY)
    u, v = U[:, 1] / rho, U[:, 2] / rho
    p = (GAMMA - 1) * (U[:, 3] - 0.5 * rho * (u ** 2 + v ** 2))
    return torch.stack([torch.log(rho), u, v, torch.log(torch.clamp(p, min=TINY))], dim=1)


def logprim_to_conserved(W):
    """(log rho, u, v, log p) -> (rho, rho u, rho v, E), positive density, pressure and energy."""
    rho, u, v, p = torch.exp(W[:, 0]), W[:, 1], W[:, 2], torch.exp(W[:, 3])
    return torch.stack([rho, rho * u, rho * v, p / (GAMMA - 1) + 0.5 * rho * (u ** 2 + v ** 2)], dim=1)


class Positive(torch.nn.Module):
    """Wraps a core model so that it steps (log rho, u, v, log p) standardised with train statistics;
    input and output stay normalised conserved variables (models.py:259-300)."""

    def __init__(self, core, u_mean, u_std, w_mean, w_std):
        super().__init__()
        self.core = core
        for name, v in (("u_mean", u_mean), ("u_std", u_std), ("w_mean", w_mean), ("w_std", w_std)):
            self.register_buffer(name, torch.as_tensor(v, dtype=torch.float32).view(1, 4, 1, 1))

    def forward(self, x):
        U = x * self.u_std + self.u_mean
        z = (conserved_to_logprim(U) - self.w_mean) / self.w_std
        W = self.core(z) * self.w_std + self.w_mean
        return (logprim_to_conserved(W) - self.u_mean) / self.u_std


def count_params(model: torch.nn.Module) -> int:
    return sum(p.numel() for p in model.parameters())


MODELS = {"resnet": ResNet, "unet": UNet, "fno": FNO}


# ------------------------------------------------------------------------------------------- training
class Normaliser:
    """Per-variable (U - mean) / std with train-split statistics (stats_<n>.json of the dataset pipeline)."""

    def __init__(self, mean, std, delta_std):
        self.mean = torch.as_tensor(mean, dtype=torch.float32).view(1, 4, 1, 1)
        self.std = torch.as_tensor(std, dtype=torch.float32).view(1, 4, 1, 1)
        self.delta_std = torch.as_tensor(delta_std, dtype=torch.float32)

    @classmethod
    def from_stats(cls, stats: dict):
        return cls(stats["mean"], stats["std"], stats["delta_std"])

    def encode(self, U):
        return (U - self.mean.to(U.device)) / self.std.to(U.device)

    def decode(self, Z):
        return Z * self.std.to(Z.device) + self.mean.to(Z.device)


def make_pairs(n_traj, n_frames, horizon):
    """Every (trajectory, start frame) whose `horizon` targets stay inside the trajectory."""
    return torch.tensor([(i, t) for i in range(n_traj) for t in range(n_frames - horizon)], dtype=torch.long)


def loss_fn(model, x, y, mode="onestep"):
    """x (B, 4, n, n); y (B, k, 4, n, n).  onestep: MSE of one step to y[:, 0]; multistep: mean MSE over k
    self-fed steps; pushforward: k-1 steps without gradient, then one trained step to y[:, -1]
    (train.py:89-111)."""
    if mode == "onestep":
        return torch.mean((model(x) - y[:, 0]) ** 2)
    if mode == "multistep":
        losses = []
        for j in range(y.shape[1]):
            x = model(x)
            losses.append(torch.mean((x - y[:, j]) ** 2))
        return torch.stack(losses).mean()
    with torch.no_grad():
        for _ in range(y.shape[1] - 1):
            x = model(x)
    return torch.mean((model(x) - y[:, -1]) ** 2)


def train(model, frames, norm: Normaliser, steps=200, batch=16, lr=1e-3, warmup=20, horizon=1, mode="onestep",
          seed=0, log=print):
    """frames: (N, F, 4, n, n) float32 tensor of train trajectories on the model's device.  AdamW with
    warm-up + cosine decay and global-norm clipping 1.0 (the reference's optimiser, train.py:209-211).
    Returns the list of losses."""
    opt = torch.optim.AdamW(model.parameters(), lr=lr, weight_decay=1e-4)
    sched = torch.optim.lr_scheduler.LambdaLR(
        opt, lambda i: min(1.0, (i + 1) / warmup) * 0.5 * (1 + math.cos(math.pi * min(1.0, i / steps))))
    pairs = make_pairs(frames.shape[0], frames.shape[1], horizon)
    g = torch.Generator().manual_seed(seed)
    hist = []
    for it in range(steps):
        sel = pairs[torch.randint(0, len(pairs), (batch,), generator=g)]
        x = norm.encode(frames[sel[:, 0], sel[:, 1]])
        y = torch.stack([norm.encode(frames[sel[:, 0], sel[:, 1] + j]) for j in range(1, horizon + 1)], dim=1)
        loss = loss_fn(model, x, y, mode)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        torch.nn.utils.clip_grad_norm_(model.parameters(), 1.0)
        opt.step()
        sched.step()
        hist.append(float(loss.detach()))
        if it % 50 == 0 or it == steps - 1:
            log(f"  step {it:5d}  loss {hist[-1]:.3e}")
    return hist


@torch.no_grad()
def rollout(model, x0, steps):
    """Self-fed rollout: (B, 4, n, n) normalised initial states -> (B, steps + 1, 4, n, n)."""
    out = [x0]
    x = x0
    for _ in range(steps):
        x = model(x)
        out.append(x)
    return torch.stack(out, dim=1)


def rollout_metrics(pred, true, std):
    """Per-frame rollout metrics of surrogate/rollout.py:54-62 plus density L1.  pred, true: physical
    (T, 4, n, n) frames 1..T of a rollout and of its truth; std (4,).  Returns a dict of
      rmse      (T, 4) RMSE per variable over the variable's std
      l1_rho    (T,)   mean |rho_pred - rho_true|
      drift     (T, 4) change of each total (mass, x/y momentum, energy) from the truth's first frame,
                       over that frame's total absolute amount (net momentum can be ~0)
      bad_cells (T,)   cells with density <= 0 or a pressure that is not > 0 (NaN counts as bad)."""
    std = torch.as_tensor(std, device=pred.device, dtype=pred.dtype)
    rmse = torch.sqrt(torch.mean((pred - true) ** 2, dim=(-2, -1))) / std
    l1_rho = torch.mean(torch.abs(pred[:, 0] - true[:, 0]), dim=(-2, -1))
    totals = pred.sum(dim=(-2, -1))
    drift = (totals - true[0].sum(dim=(-2, -1))) / true[0].abs().sum(dim=(-2, -1))
    rho, mx, my, E = pred[:, 0], pred[:, 1], pred[:, 2], pred[:, 3]
    p = (GAMMA - 1.0) * (E - 0.5 * (mx * mx + my * my) / rho)
    bad = ((rho <= 0) | ~(p > 0)).sum(dim=(-2, -1))
    return {"rmse": rmse, "l1_rho": l1_rho, "drift": drift, "bad_cells": bad}

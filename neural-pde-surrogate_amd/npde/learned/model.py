"""Conservative learned flux correction for the coarse solver.

A stack of 3x3 periodic convolutions (4 hidden layers of 32 channels with tanh-GELU, then a linear
head; 32,392 parameters) maps the coarse state to 8 per-face flux increments: 4 conserved components
on the face left of each cell and 4 on the face below it.  Each increment is out_scale * tanh(raw)
times the face's natural flux scale (rho a, rho a^2, rho a^2, rho a^3 with the smaller density and
the larger sound speed of the two cells sharing the face), so a correction stays a bounded fraction of
the flux it corrects, also in rarefied, low-pressure regions where a fixed-size increment would drive
the state non-physical.  The increments are added to the solver's
face fluxes (HLLC + van Leer) inside the stage kernels, so the update stays in flux form: mass,
momentum and energy are conserved to rounding whatever the network outputs.  The head is zero
initialised, so an untrained corrector reproduces the plain solver exactly (analog: the zero-init
last layer of the reference's surrogates, surrogate/models.py:66-69).

Inputs are 8 features per cell (conserved and primitive variables and sound speed, shifted and
scaled to order one) padded to 16 channels for the MFMA kernels.
"""
from __future__ import annotations

import math

import torch

from .conv import FluxCNN, flux_cnn_infer, pack_stack

GAMMA = 1.4
IN_FEATURES = 8
# (offset, scale) per feature: rho, rho*u, rho*v, E, u, v, p, a
FEATURE_NORM = [(1.0, 1.0), (0.0, 1.0), (0.0, 1.0), (2.5, 2.5), (0.0, 1.0), (0.0, 1.0), (1.0, 1.0), (1.2, 0.6)]


def primitives(U: torch.Tensor):
    d, m, n, e = U[:, 0], U[:, 1], U[:, 2], U[:, 3]
    u, v = m / d, n / d
    p = (GAMMA - 1.0) * (e - 0.5 * d * (u * u + v * v))
    a = torch.sqrt(torch.clamp(GAMMA * p / d, min=0.0))
    return d, u, v, p, a


def features(U: torch.Tensor) -> torch.Tensor:
    """(B, 4, H, W) float64 conserved -> (B, H, W, 16) float32 normalised features (channels 8-15 zero)."""
    d, u, v, p, a = primitives(U)
    # scalar constants only (no host-to-device copies), so the function can be graph-captured
    vals = [d, U[:, 1], U[:, 2], U[:, 3], u, v, p, a]
    f = torch.stack([(x - o) / s for x, (o, s) in zip(vals, FEATURE_NORM)], dim=-1).to(torch.float32)
    return torch.cat([f, torch.zeros_like(f)], dim=-1).contiguous()


def face_scales(U: torch.Tensor):
    """Flux scales (B, 4, H, W) of the face left of each cell (x) and below it (y), periodic neighbours."""
    d, _, _, _, a = primitives(U)
    out = []
    for dim in (-1, -2):
        df = torch.minimum(d, d.roll(1, dims=dim))
        af = torch.maximum(a, a.roll(1, dims=dim))
        out.append(torch.stack([df * af, df * af * af, df * af * af, df * af * af * af], dim=1))
    return out


class FluxCorrector(torch.nn.Module):
    def __init__(self, hidden: int = 32, hidden_layers: int = 4, out_scale: float = 0.02, seed: int = 0):
        super().__init__()
        chans = [IN_FEATURES] + [hidden] * hidden_layers + [8]
        g = torch.Generator().manual_seed(seed)
        self.weights = torch.nn.ParameterList()
        self.biases = torch.nn.ParameterList()
        for k in range(len(chans) - 1):
            cin, cout = chans[k], chans[k + 1]
            last = k == len(chans) - 2
            std = 0.0 if last else math.sqrt(2.0 / (9 * cin))
            w = torch.randn(cout, cin, 3, 3, generator=g) * std
            self.weights.append(torch.nn.Parameter(w))
            self.biases.append(torch.nn.Parameter(torch.zeros(cout)))
        self.out_scale = out_scale

    def params(self):
        out = []
        for w, b in zip(self.weights, self.biases):
            out += [w, b]
        return out

    def num_parameters(self) -> int:
        return sum(p.numel() for p in self.parameters())

    def forward(self, U: torch.Tensor):
        """(B, 4, H, W) float64 state -> (corr_x, corr_y), each (B, 4, H, W) float64."""
        # the network's input is not differentiated: gradients reach it through the flux increments only
        Ud = U.detach()
        out = FluxCNN.apply(features(Ud), *self.params())  # (B, H, W, 8) float32
        out = (torch.tanh(out) * self.out_scale).to(torch.float64)
        sx, sy = face_scales(Ud)
        cx = (out[..., 0:4].permute(0, 3, 1, 2) * sx).contiguous()
        cy = (out[..., 4:8].permute(0, 3, 1, 2) * sy).contiguous()
        return cx, cy

    def increments_into(self, U: torch.Tensor, CX: torch.Tensor, CY: torch.Tensor, keep: torch.Tensor, packed):
        """Inference: the increments of forward(U) written into CX, CY (B, 4, H, W) float64, zero for
        instances whose keep flag is False.  `packed` = pack_stack(self.params()), computed once; no
        tensor is allocated outside torch's allocator and nothing synchronises the host, so the call
        can be captured in a graph."""
        out = flux_cnn_infer(features(U), packed)
        out = (torch.tanh(out) * self.out_scale).to(torch.float64)
        sx, sy = face_scales(U)
        k = keep.view(-1, 1, 1, 1)
        zero = torch.zeros((), dtype=torch.float64, device=U.device)
        CX.copy_(torch.where(k, out[..., 0:4].permute(0, 3, 1, 2) * sx, zero))
        CY.copy_(torch.where(k, out[..., 4:8].permute(0, 3, 1, 2) * sy, zero))

    def state(self) -> dict:
        return {f"w{k}": w.detach().cpu() for k, w in enumerate(self.weights)} | \
               {f"b{k}": b.detach().cpu() for k, b in enumerate(self.biases)} | \
               {"out_scale": torch.tensor(self.out_scale)}

    def load(self, st: dict):
        for k in range(len(self.weights)):
            self.weights[k].data.copy_(torch.as_tensor(st[f"w{k}"]))
            self.biases[k].data.copy_(torch.as_tensor(st[f"b{k}"]))
        self.out_scale = float(st["out_scale"])


TRAINED_WEIGHTS = __import__("pathlib").Path(__file__).resolve().parent / "weights" / "flux_corrector_riemann64.npz"


def load_trained(path=TRAINED_WEIGHTS, device="cuda") -> FluxCorrector:
    """The corrector trained on the random-Riemann family at 64^2 (weights stored as a plain .npz of
    float32 arrays, read without pickle)."""
    import numpy as np
    st = dict(np.load(path, allow_pickle=False))
    m = FluxCorrector()
    m.load({k: torch.from_numpy(v) for k, v in st.items()})
    return m.to(device)

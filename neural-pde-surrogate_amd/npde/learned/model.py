"""Conservative learned flux correction for the coarse solver.

A stack of 3x3 periodic convolutions (4 hidden layers of 32 channels with tanh-GELU, then a linear
head; 32,392 parameters) maps the coarse state to 8 per-face factors: 4 conserved components on the
face left of each cell and 4 on the face below it, each out_scale * tanh(raw).  The solver adds to each
face flux its factor times the face's natural flux scale (rho a, rho a^2, rho a^2, rho a^3 with the
smaller density and the larger sound speed of the two cells sharing the face), evaluated on every RK
stage's input state (csrc/euler_numerics.hpp corr_scale; npde.learned.torch_solver.face_scale): a
correction stays a bounded fraction of the local flux as the flow evolves between network calls, also
in rarefied, low-pressure regions where a fixed-size increment would drive the state non-physical.  The
increments enter the face fluxes (HLLC + van Leer) inside the stage kernels, so the update stays in flux
form: mass, momentum and energy are conserved to rounding whatever the network outputs.  The head is
zero initialised, so an untrained corrector reproduces the plain solver exactly (analog: the zero-init
last layer of the reference's surrogates, surrogate/models.py:66-69).

Inputs are 8 features per cell (conserved and primitive variables and sound speed, shifted and
scaled to order one); the per-layer training kernels take them padded to 16 channels, the fused
inference kernel (npde_flux_net) takes the 8 and computes them itself from the state.
"""
from __future__ import annotations

import math

import torch

from .conv import FluxCNN

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


class FluxCorrector(torch.nn.Module):
    def __init__(self, hidden: int = 32, hidden_layers: int = 4, out_scale: float = 0.02, seed: int = 0,
                 calls_per_frame: int = 1):
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
        # network evaluations per frame interval: the factors are recomputed at calls_per_frame equally
        # spaced times inside each interval (the solver lands on them as on frame times)
        self.calls_per_frame = int(calls_per_frame)

    def params(self):
        out = []
        for w, b in zip(self.weights, self.biases):
            out += [w, b]
        return out

    def num_parameters(self) -> int:
        return sum(p.numel() for p in self.parameters())

    def forward(self, U: torch.Tensor):
        """(B, 4, H, W) float64 state -> (corr_x, corr_y), each (B, 4, H, W) float64 dimensionless factors
        of the face left of / below each cell (the solver multiplies them by the face flux scales)."""
        # the network's input is not differentiated: gradients reach it through the flux increments only
        out = FluxCNN.apply(features(U.detach()), *self.params())  # (B, H, W, 8) float32
        out = (torch.tanh(out) * self.out_scale).to(torch.float64)
        return out[..., 0:4].permute(0, 3, 1, 2).contiguous(), out[..., 4:8].permute(0, 3, 1, 2).contiguous()

    def pack_fused(self):
        """The weights and biases in the operand order of the fused inference kernel (npde_set_flux_network):
        (fragments, biases, out_scale) with fragments an int16 tensor of 69 x 64 x 8 bfloat16 bit patterns
        (round to nearest even) and biases 136 float32, on the model's device.  A fragment holds, for each of
        the 64 lanes l (m = l % 16, g = l / 16), the 8 k-values of output channel 16 t + m: input channels
        8 g .. 8 g + 7 of tap b for the 32-channel layers (k-block b = tap), the 8 features of tap 4 b + g for
        the first layer (zero past the ninth tap); tap = ky * 3 + kx of the 3 x 3 cross-correlation."""
        def first(w):  # (32, 8, 3, 3) -> (3, 2, 4, 16, 8)
            wt = w.reshape(32, 8, 9)
            out = w.new_zeros(3, 2, 4, 16, 8)
            for b in range(3):
                for g in range(4):
                    if 4 * b + g < 9:
                        out[b, :, g] = wt[:, :, 4 * b + g].reshape(2, 16, 8)
            return out

        def hidden(w, cout_p):  # (cout, 32, 3, 3) -> (9, cout_p / 16, 4, 16, 8)
            cout = w.shape[0]
            wt = w.new_zeros(cout_p, 32, 9)
            wt[:cout] = w.reshape(cout, 32, 9)
            return wt.permute(2, 0, 1).reshape(9, cout_p // 16, 16, 4, 8).permute(0, 1, 3, 2, 4)

        ws = [w.detach().float() for w in self.weights]
        if [tuple(w.shape) for w in ws] != [(32, 8, 3, 3)] + [(32, 32, 3, 3)] * 3 + [(8, 32, 3, 3)]:
            raise ValueError("pack_fused needs the 8 -> 4 x 32 -> 8 stack")
        frags = torch.cat([first(ws[0]).reshape(-1)] + [hidden(w, 32).reshape(-1) for w in ws[1:4]]
                          + [hidden(ws[4], 16).reshape(-1)])
        assert frags.numel() == 69 * 64 * 8
        bits = frags.to(torch.bfloat16).view(torch.int16).contiguous()
        biases = torch.cat([b.detach().float() for b in self.biases]).contiguous()
        return bits, biases, float(self.out_scale)

    def state(self) -> dict:
        return {f"w{k}": w.detach().cpu() for k, w in enumerate(self.weights)} | \
               {f"b{k}": b.detach().cpu() for k, b in enumerate(self.biases)} | \
               {"out_scale": torch.tensor(self.out_scale), "calls_per_frame": torch.tensor(self.calls_per_frame)}

    def load(self, st: dict):
        for k in range(len(self.weights)):
            self.weights[k].data.copy_(torch.as_tensor(st[f"w{k}"]))
            self.biases[k].data.copy_(torch.as_tensor(st[f"b{k}"]))
        self.out_scale = float(st["out_scale"])
        self.calls_per_frame = int(st["calls_per_frame"]) if "calls_per_frame" in st else 1


TRAINED_WEIGHTS = __import__("pathlib").Path(__file__).resolve().parent / "weights" / "flux_corrector_riemann64.npz"


def load_trained(path=TRAINED_WEIGHTS, device="cuda") -> FluxCorrector:
    """The corrector trained on the random-Riemann family at 64^2 (weights stored as a plain .npz of
    float32 arrays, read without pickle)."""
    import numpy as np
    st = dict(np.load(path, allow_pickle=False))
    m = FluxCorrector()
    m.load({k: torch.from_numpy(v) for k, v in st.items()})
    return m.to(device)

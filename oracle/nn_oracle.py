"""CPU binary64 restatement of the flux-correction network -- TEST INFRASTRUCTURE ONLY (imported by
tests/ as the checker of the BF16-MFMA kernels, never by the product).

Same stack as npde.learned (3x3 circular convolutions, tanh-approximate GELU on hidden layers, linear
head), evaluated in float64 with torch.nn.functional.conv2d, with the operands rounded to bfloat16 at
the points where the GPU kernels round them (layer inputs and weights).  Gradients come from torch's
autograd through that float64 graph (the roundings pass gradients straight through).
"""
from __future__ import annotations

import math

import torch
import torch.nn.functional as F

C0 = math.sqrt(2.0 / math.pi)
C1 = 0.044715


def gelu_tanh(x):
    return 0.5 * x * (1.0 + torch.tanh(C0 * (x + C1 * x * x * x)))


def bf16(x):
    """round to bfloat16 and back, gradient passed straight through"""
    return x + (x.to(torch.bfloat16).to(x.dtype) - x).detach()


def forward(x_nhwc, weights, biases, round_bf16=True):
    """x_nhwc: (B, H, W, C >= Cin) -> (B, H, W, Cout_last), all float64 on the CPU."""
    h = x_nhwc.permute(0, 3, 1, 2)[:, :weights[0].shape[1]]
    L = len(weights)
    for k in range(L):
        hin = bf16(h) if round_bf16 else h
        w = bf16(weights[k]) if round_bf16 else weights[k]
        z = F.conv2d(F.pad(hin, (1, 1, 1, 1), mode="circular"), w) + biases[k].view(1, -1, 1, 1)
        h = gelu_tanh(z) if k < L - 1 else z
    return h.permute(0, 2, 3, 1)

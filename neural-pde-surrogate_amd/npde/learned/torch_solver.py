"""Differentiable binary64 restatement of the FV step, batched, for training the flux correction.

The learned coarse solver runs on the HIP stage kernels (flux increments added inside the kernel).
Training needs gradients of a rollout loss through solver steps with respect to those increments;
this module provides the same step as torch operations so autograd can supply them.  Every
element's expression keeps the reference's operation order (state2d.hpp, limiters.hpp,
riemann2d.hpp, solver2d.hpp:86-180), so on the same input it reproduces the kernels' results bit for
bit (checked in tests/test_learned.py against the CPU oracle and the HIP kernels).  Operations are
vectorised over the four variables and over both face directions (x and y faces solved in one
Riemann call), which keeps the number of launched ops per step small.  Unselected branches use
harmless denominators so their gradients stay finite; the selected values are unchanged.

State: (B, 4, ny, nx) float64 on a periodic grid.  Corrections: (B, 4, ny, nx) each, dimensionless
factors for the face left of a cell (cx) and the face below it (cy); in both RK stages each face flux
gets its factor times the face's flux scale (face_scale, from the stage's input state), as the HIP stage
kernels do (csrc/euler_numerics.hpp corr_scale).
"""
from __future__ import annotations

import torch

GAMMA = 1.4
G1 = GAMMA - 1.0


class _Sqrt(torch.autograd.Function):
    """Correctly rounded square root (torch's CPU kernel is a vectorised approximation; numpy's is
    the IEEE one); derivative 1 / (2 sqrt x)."""

    @staticmethod
    def forward(ctx, x):
        if x.is_cuda:
            y = torch.sqrt(x)
        else:
            import numpy as np
            y = torch.from_numpy(np.sqrt(x.detach().numpy()))
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, g):
        (y,) = ctx.saved_tensors
        return g * (0.5 / y)


def _sqrt(x):
    return _Sqrt.apply(x)


def _c(v, like):
    """a constant as a 0-d tensor on `like`'s device: torch's GPU kernels turn division by a Python
    scalar into multiplication by its reciprocal, which is not the IEEE quotient."""
    return torch.tensor(v, dtype=like.dtype, device=like.device)


def _tmin(a, b):
    return torch.where(b < a, b, a)


def _tmax(a, b):
    return torch.where(a < b, b, a)


def prim(d, m, n, e):
    u = m / d
    v = n / d
    return d, u, v, G1 * (e - (0.5 * d) * (u * u + v * v))


def limit(kind: str, a, b):
    if kind == "none":
        return torch.zeros_like(a)
    if kind == "unlimited":
        return 0.5 * (a + b)
    ab = a * b
    if kind == "minmod":
        return torch.where(ab <= 0.0, torch.zeros_like(a), torch.where(a.abs() < b.abs(), a, b))
    den = torch.where(ab <= 0.0, torch.ones_like(a), a + b)
    return torch.where(ab <= 0.0, torch.zeros_like(a), 2.0 * a * b / den)


def sides(E):
    """E (..., 4, H, W) face values of (rho, u, v, p) -> to_conserved, to_primitive, sound speed,
    as a dict of (..., H, W) tensors in the x frame (mn = rho*u, un = u)."""
    d, u, v, p = E.unbind(-3)
    kin = (0.5 * d) * (u * u + v * v)
    m, n, e = d * u, d * v, p / _c(G1, p) + kin
    _, u2, v2, p2 = prim(d, m, n, e)
    return {"d": d, "mn": m, "mt": n, "e": e, "un": u2, "ut": v2, "p": p2, "a": _sqrt((GAMMA * p2) / d)}


def _phys(s):
    return s["mn"], s["mn"] * s["un"] + s["p"], s["mt"] * s["un"], s["un"] * (s["e"] + s["p"])


def hllc(A, B):
    sl = _tmin(A["un"] - A["a"], B["un"] - B["a"])
    sr = _tmax(A["un"] + A["a"], B["un"] + B["a"])
    num = ((B["p"] - A["p"]) + (A["d"] * A["un"]) * (sl - A["un"])) - (B["d"] * B["un"]) * (sr - B["un"])
    den = A["d"] * (sl - A["un"]) - B["d"] * (sr - B["un"])
    ss = num / den
    left = (0.0 <= sl) | (0.0 <= ss)
    star = torch.where(left, ~(0.0 <= sl), 0.0 <= sr)
    X = {k: torch.where(left, A[k], B[k]) for k in A}
    S = torch.where(left, sl, sr)
    F = _phys(X)
    q = X["d"] * (S - X["un"])
    gap = torch.where(star, S - ss, torch.ones_like(S))
    qs = torch.where(star, q, torch.ones_like(q))
    f = q / gap
    st = (f, f * ss, f * X["ut"], f * (X["e"] / X["d"] + (ss - X["un"]) * (ss + X["p"] / qs)))
    U = (X["d"], X["mn"], X["mt"], X["e"])
    return tuple(torch.where(star, F[k] + S * (st[k] - U[k]), F[k]) for k in range(4))


def rusanov(A, B):
    s = _tmax(A["un"].abs() + A["a"], B["un"].abs() + B["a"])
    FA, FB = _phys(A), _phys(B)
    h = 0.5 * s
    UA = (A["d"], A["mn"], A["mt"], A["e"])
    UB = (B["d"], B["mn"], B["mt"], B["e"])
    return tuple(0.5 * (FA[k] + FB[k]) - h * (UB[k] - UA[k]) for k in range(4))


_YSWAP = {"d": "d", "mn": "mt", "mt": "mn", "e": "e", "un": "ut", "ut": "un", "p": "p", "a": "a"}


CORR_RHO = 0.05  # densities below this fade the corrections out (csrc/euler_numerics.hpp kCorrRho)


def face_scale(dl, al, dr, ar):
    """(B, H, W) densities and sound speeds of the cells left of / below (l) and right of / above (r) each
    face -> (B, 4, H, W) flux scales r a, r a^2, r a^2, r a^3 with a the larger sound speed and
    r = d^2 / (d + CORR_RHO) for d the smaller density (the kernels' corr_scale, same operation order)."""
    d, a = _tmin(dl, dr), _tmax(al, ar)
    r = (d * d) / (d + CORR_RHO)
    s1 = r * a
    s2 = s1 * a
    return torch.stack([s1, s2, s2, s2 * a], dim=1)


def rhs(U, dx, dy, scheme="hllc", lim="vanleer", cx_corr=None, cy_corr=None):
    """L(U) for a periodic batch (B, 4, ny, nx); x faces first, then y (solver2d.hpp:129-180)."""
    riem = hllc if scheme == "hllc" else rusanov
    d, u, v, p = prim(U[:, 0], U[:, 1], U[:, 2], U[:, 3])
    W = torch.stack([d, u, v, p], dim=1)                                   # (B, 4, H, W)
    W2 = torch.stack([W, W])                                               # dir x, dir y
    Wm = torch.stack([torch.roll(W, 1, -1), torch.roll(W, 1, -2)])         # cell i-1 / j-1
    Wp = torch.stack([torch.roll(W, -1, -1), torch.roll(W, -1, -2)])       # cell i+1 / j+1
    dW = limit(lim, W2 - Wm, Wp - W2)
    h = torch.tensor([-0.5, 0.5], dtype=U.dtype, device=U.device).view(2, 1, 1, 1, 1, 1)
    S = sides(W2.unsqueeze(0) + h * dW.unsqueeze(0))                       # (sign, dir, B, H, W) fields
    # face left of / below each cell: A = right/top side of the previous cell, B = left/bottom side here
    A, Bs = {}, {}
    for k in S:
        ax = torch.roll(S[k][1, 0], 1, -1)
        ay = torch.roll(S[_YSWAP[k]][1, 1], 1, -2)
        A[k] = torch.stack([ax, ay])
        Bs[k] = torch.stack([S[k][0, 0], S[_YSWAP[k]][0, 1]])
    f = riem(A, Bs)                                                        # normal frame, (dir, B, H, W)
    Fl = torch.stack([f[0][0], f[1][0], f[2][0], f[3][0]], dim=1)          # x faces: (rho, mx, my, E)
    Gb = torch.stack([f[0][1], f[2][1], f[1][1], f[3][1]], dim=1)          # y faces back to (rho, mx, my, E)
    if cx_corr is not None or cy_corr is not None:
        a = _sqrt((GAMMA * p) / d)
        if cx_corr is not None:
            Fl = Fl + cx_corr * face_scale(torch.roll(d, 1, -1), torch.roll(a, 1, -1), d, a)
        if cy_corr is not None:
            Gb = Gb + cy_corr * face_scale(torch.roll(d, 1, -2), torch.roll(a, 1, -2), d, a)
    cx, cy = 1.0 / dx, 1.0 / dy
    Lx = (0.0 + cx * Fl) - cx * torch.roll(Fl, -1, -1)
    return (Lx + cy * Gb) - cy * torch.roll(Gb, -1, -2)


def wavespeed_dt(U, dx, dy, cfl=0.5):
    """compute_dt (solver2d.hpp:67-83) per instance, no gradient: returns (B,) float64."""
    with torch.no_grad():
        d, u, v, p = prim(U[:, 0], U[:, 1], U[:, 2], U[:, 3])
        a = _sqrt((GAMMA * p) / d)
        sx = torch.clamp((u.abs() + a).flatten(1).amax(1), min=0.0)
        sy = torch.clamp((v.abs() + a).flatten(1).amax(1), min=0.0)
        return torch.full_like(sx, cfl) / (sx / _c(dx, sx) + sy / _c(dy, sy))


def step(U, dt, dx, dy, scheme="hllc", lim="vanleer", cx_corr=None, cy_corr=None):
    """SSP-RK2 (solver2d.hpp:86-103); dt is a (B,) tensor."""
    dtb = dt.view(-1, 1, 1, 1)
    U1 = U + dtb * rhs(U, dx, dy, scheme, lim, cx_corr, cy_corr)
    return 0.5 * (U + U1 + dtb * rhs(U1, dx, dy, scheme, lim, cx_corr, cy_corr))


def advance(U, t, t_next, dx, dy, cfl=0.5, scheme="hllc", lim="vanleer", cx_corr=None, cy_corr=None,
            max_steps=100000):
    """Step every instance until t >= t_next with landing dt = min(dt_cfl, t_next - t) (main2d.cpp:237-242).
    t: (B,) float64 tensor.  Instances that have arrived are carried unchanged."""
    steps = 0
    while bool((t < t_next).any()) and steps < max_steps:
        dtc = wavespeed_dt(U, dx, dy, cfl)
        gap = t_next - t
        dt = torch.where(gap < dtc, gap, dtc)
        live = t < t_next
        dt = torch.where(live, dt, torch.zeros_like(dt))
        Un = step(U, dt, dx, dy, scheme, lim, cx_corr, cy_corr)
        U = torch.where(live.view(-1, 1, 1, 1), Un, U)
        t = torch.where(live, t + dt, t)
        steps += 1
    return U, t, steps

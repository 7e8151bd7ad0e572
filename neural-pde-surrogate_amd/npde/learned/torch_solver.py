"""Differentiable binary64 restatement of the FV step, batched, for training the flux correction.

The learned coarse solver runs on the HIP stage kernels (flux increments added inside the kernel).
Training needs gradients of a rollout loss through solver steps with respect to those increments;
this module provides the same step as torch operations so autograd can supply them.  Every
expression keeps the reference's operation order (state2d.hpp, limiters.hpp, riemann2d.hpp,
solver2d.hpp:86-180), so on the same input it reproduces the kernels' results bit for bit (checked in
tests/test_learned.py against the CPU oracle).  Unselected branches use harmless denominators so that
their gradients stay finite; the selected values are unchanged.

State: (B, 4, ny, nx) float64 on a periodic grid.  Corrections: (B, 4, ny, nx) each, cx on the face
left of a cell and cy on the face below it, added to the face fluxes in both RK stages.
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


def side(w, dw, h):
    """face-side state at W + h*dW: to_conserved then to_primitive, plus sound speed (x frame)."""
    d = w[0] + h * dw[0]
    u = w[1] + h * dw[1]
    v = w[2] + h * dw[2]
    p = w[3] + h * dw[3]
    kin = (0.5 * d) * (u * u + v * v)
    m, n, e = d * u, d * v, p / _c(G1, p) + kin
    _, u2, v2, p2 = prim(d, m, n, e)
    a = _sqrt((GAMMA * p2) / d)
    return {"d": d, "mn": m, "mt": n, "e": e, "un": u2, "ut": v2, "p": p2, "a": a}


def y_frame(s):
    return {"d": s["d"], "mn": s["mt"], "mt": s["mn"], "e": s["e"], "un": s["ut"], "ut": s["un"], "p": s["p"],
            "a": s["a"]}


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


def _shift(x, k, dim):
    """value of the cell k places back along dim (periodic): out[i] = x[i - k]"""
    return torch.roll(x, shifts=k, dims=dim)


def rhs(U, dx, dy, scheme="hllc", lim="vanleer", cx_corr=None, cy_corr=None):
    """L(U) for a periodic batch (B, 4, ny, nx); x faces first, then y (solver2d.hpp:129-180)."""
    riem = hllc if scheme == "hllc" else rusanov
    w = prim(U[:, 0], U[:, 1], U[:, 2], U[:, 3])
    cx, cy = 1.0 / dx, 1.0 / dy
    # x direction (dim -1)
    wm = [_shift(q, 1, -1) for q in w]
    wp = [_shift(q, -1, -1) for q in w]
    dws = [limit(lim, w[k] - wm[k], wp[k] - w[k]) for k in range(4)]
    L_ = side(w, dws, -0.5)
    R_ = side(w, dws, +0.5)
    Rprev = {k: _shift(v, 1, -1) for k, v in R_.items()}
    f = riem(Rprev, L_)                                   # face left of each cell: (d, mn=mx, mt=my, e)
    fl = [f[0], f[1], f[2], f[3]]
    if cx_corr is not None:
        fl = [fl[q] + cx_corr[:, q] for q in range(4)]
    fr = [_shift(q, -1, -1) for q in fl]
    Lx = [(0.0 + cx * fl[q]) - cx * fr[q] for q in range(4)]
    # y direction (dim -2)
    wm = [_shift(q, 1, -2) for q in w]
    wp = [_shift(q, -1, -2) for q in w]
    dws = [limit(lim, w[k] - wm[k], wp[k] - w[k]) for k in range(4)]
    Bo = y_frame(side(w, dws, -0.5))
    To = y_frame(side(w, dws, +0.5))
    Tprev = {k: _shift(v, 1, -2) for k, v in To.items()}
    g = riem(Tprev, Bo)                                   # y frame: (d, mn=my, mt=mx, e)
    gb = [g[0], g[2], g[1], g[3]]                         # back to (rho, mx, my, E): face below each cell
    if cy_corr is not None:
        gb = [gb[q] + cy_corr[:, q] for q in range(4)]
    gt = [_shift(q, -1, -2) for q in gb]
    return torch.stack([(Lx[q] + cy * gb[q]) - cy * gt[q] for q in range(4)], dim=1)


def wavespeed_dt(U, dx, dy, cfl=0.5):
    """compute_dt (solver2d.hpp:67-83) per instance, no gradient: returns (B,) float64."""
    with torch.no_grad():
        d, u, v, p = prim(U[:, 0], U[:, 1], U[:, 2], U[:, 3])
        a = _sqrt((GAMMA * p) / d)
        sx = (u.abs() + a).flatten(1).amax(1)
        sy = (v.abs() + a).flatten(1).amax(1)
        sx = torch.clamp(sx, min=0.0)
        sy = torch.clamp(sy, min=0.0)
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

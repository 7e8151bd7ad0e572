"""Seeded initial-condition families of the benchmark workloads.

Every sample is a pure function of (family, index): the generator is numpy's PCG64 seeded with
[crc32(family), index], the reference's seeding scheme (data/families.py:16-19), so any instance of
a batch can be regenerated alone.  The distributions are the benchmark's own (they differ from the
reference's dataset families, whose densities and pressures are log-uniform):

  riemann  four constant quadrants split at a corner (cx, cy) ~ U[0.3, 0.7]^2, each quadrant
           rho ~ U[0.1, 2], u, v ~ U[-1, 1], p ~ U[0.1, 2]; periodic unit box
  blast    background rho = 1, u = v = 0, p = 0.1; 1-3 blasts at uniform centres with radius
           ~ U[0.03, 0.1] and p ~ U[5, 50]; 0-3 density bubbles with radius ~ U[0.05, 0.15] and
           rho ~ U[0.1, 5]; discs smoothed with tanh edges of width 0.005 on the periodic distance
           (reference: data/families.py:70-79)
  kh       two tanh shear layers at y = 1/4, 3/4 with smoothing length 0.025, p = 2.5, density ratio
           ~ U[1.5, 3], shear du ~ U[0.5, 1.5], v = A sin(4 pi x) with A ~ U[0.005, 0.02]

Each function returns (U, t_end, params) with U the conserved planes (4, n, n).
"""
from __future__ import annotations

import zlib

import numpy as np

from .solver import to_conserved


def rng_for(family: str, index: int) -> np.random.Generator:
    return np.random.default_rng([zlib.crc32(family.encode()), int(index)])


def cell_centres(n: int) -> np.ndarray:
    return (np.arange(n) + 0.5) / n


def riemann(index: int, n: int, family: str = "riemann_b1"):
    rng = rng_for(family, index)
    cx, cy = rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7)
    # quadrants: 0 bottom-left, 1 bottom-right, 2 top-left, 3 top-right; columns rho, u, v, p
    states = np.array([[rng.uniform(0.1, 2.0), rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0),
                        rng.uniform(0.1, 2.0)] for _ in range(4)])
    # the quadrant of cell (i, j) is (x_i > cx) + 2 (y_j > cy); the centres increase, so each quadrant is
    # a block and its conserved state (to_conserved of its primitives, the same operations per cell) is
    # written once per block
    x = cell_centres(n)
    i0, j0 = int(np.count_nonzero(~(x > cx))), int(np.count_nonzero(~(x > cy)))
    cons = to_conserved(*(states[:, q] for q in range(4)))  # (4 variables, 4 quadrants)
    U = np.empty((4, n, n))
    for q, (rows, cols) in enumerate(((slice(0, j0), slice(0, i0)), (slice(0, j0), slice(i0, n)),
                                      (slice(j0, n), slice(0, i0)), (slice(j0, n), slice(i0, n)))):
        U[:, rows, cols] = cons[:, q, None, None]
    return U, 0.5, {"cx": cx, "cy": cy, "states": states}


def _disc(r, radius, width=0.005):
    return 0.5 * (1.0 - np.tanh((r - radius) / width))


# numpy's tanh is exactly 1.0 from 19 on, so the smoothed disc is exactly 0 -- and the update
# A + (v - A) * 0 leaves A unchanged -- wherever the periodic distance exceeds radius + 20 widths
_DISC_REACH = 20 * 0.005


def _add_disc(A, x, cx, cy, radius, value):
    """A + (value - A) * _disc(d, radius) with d the periodic distance of each cell centre from (cx, cy):
    d = sqrt(dx*dx + dy*dy), dx = (x - cx + 0.5) % 1 - 0.5 and likewise dy (reference:
    data/families.py:70-79).  Evaluated on the rows and columns within radius + _DISC_REACH of the
    centre along each axis and left unchanged elsewhere, where the disc is exactly 0."""
    dx = (x - cx + 0.5) % 1.0 - 0.5
    dy = (x - cy + 0.5) % 1.0 - 0.5
    cols = np.nonzero(np.abs(dx) < radius + _DISC_REACH)[0]
    rows = np.nonzero(np.abs(dy) < radius + _DISC_REACH)[0]
    if cols.size == 0 or rows.size == 0:
        return A
    ddx, ddy = dx[cols][None, :], dy[rows][:, None]
    s = _disc(np.sqrt(ddx * ddx + ddy * ddy), radius)
    block = np.ix_(rows, cols)
    A[block] = A[block] + (value - A[block]) * s
    return A


BLAST_SIZES = (128, 256, 512, 1024)
BLAST_WEIGHTS = (4, 3, 2, 1)


def blast_size(index: int, family: str = "blast_b2") -> int:
    rng = rng_for(family + "_size", index)
    w = np.array(BLAST_WEIGHTS, dtype=float)
    return int(rng.choice(BLAST_SIZES, p=w / w.sum()))


def blast(index: int, n: int | None = None, family: str = "blast_b2"):
    n = blast_size(index, family) if n is None else n
    rng = rng_for(family, index)
    x = cell_centres(n)
    rho = np.ones((n, n))
    p = np.full((n, n), 0.1)
    blasts, bubbles = [], []
    for _ in range(int(rng.integers(1, 4))):
        cx, cy = rng.uniform(0, 1), rng.uniform(0, 1)
        r, pb = rng.uniform(0.03, 0.1), rng.uniform(5.0, 50.0)
        p = _add_disc(p, x, cx, cy, r, pb)
        blasts.append((cx, cy, r, pb))
    for _ in range(int(rng.integers(0, 4))):
        cx, cy = rng.uniform(0, 1), rng.uniform(0, 1)
        r, rb = rng.uniform(0.05, 0.15), rng.uniform(0.1, 5.0)
        rho = _add_disc(rho, x, cx, cy, r, rb)
        bubbles.append((cx, cy, r, rb))
    z = np.zeros((n, n))
    return to_conserved(rho, z, z, p), 0.2, {"n": n, "blasts": blasts, "bubbles": bubbles}


def kh(index: int, n: int, family: str = "kh_b4"):
    rng = rng_for(family, index)
    ratio, du, amp = rng.uniform(1.5, 3.0), rng.uniform(0.5, 1.5), rng.uniform(0.005, 0.02)
    x = cell_centres(n)
    X, Y = np.meshgrid(x, x)
    L = 0.025
    band = 0.5 * (np.tanh((Y - 0.25) / L) - np.tanh((Y - 0.75) / L))  # ~1 in the middle strip
    rho = 1.0 + (ratio - 1.0) * band
    u = 0.5 * du - du * band
    v = amp * np.sin(4.0 * np.pi * X)
    p = np.full((n, n), 2.5)
    t_start = rng.uniform(0.5, 1.5)
    return to_conserved(rho, u, v, p), 2.0, {"ratio": ratio, "du": du, "amp": amp, "t_start": t_start}


# ---------------------------------------------------------------------------------------------------
# The reference's own dataset families (data/families.py), for datasets interchangeable with the
# reference's packed files.  Same seeding (rng_for), same draws in the same order and the same per-cell
# formulas as data/families.py:35-171 (densities, pressures and ratios log-uniform), with the cell centres
# of data/ic.py:15-17 (lo + (i + 0.5) * dx).  Each returns (U, t_start, t_end, params) with U the conserved
# planes (4, n, n) of the primitive fields the reference writes into its initial-condition file; its
# driver converts them with to_conserved on reading (io_hdf5.hpp:175-206).  The recorded window is
# t_start .. t_end with 50 frames (data/generate.py:60-75).

def _ref_centres(n: int) -> np.ndarray:
    return 0.0 + (np.arange(n) + 0.5) * ((1.0 - 0.0) / n)


def _log_uniform(rng: np.random.Generator, lo: float, hi: float) -> float:
    return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def _ref_disc(X, Y, cx, cy, radius, width=0.005):
    """Smoothed indicator of the disc of the given radius around (cx, cy) on the periodic unit box."""
    dx = (X - cx + 0.5) % 1.0 - 0.5
    dy = (Y - cy + 0.5) % 1.0 - 0.5
    r = np.sqrt(dx ** 2 + dy ** 2)
    return 0.5 * (1.0 - np.tanh((r - radius) / width))


def ref_riemann(index: int, n: int):
    """data/families.py:35-51: corner (xs, ys) ~ U[0.3, 0.7]^2, per quadrant (rho, u, v, p) with rho, p
    log-uniform in [0.1, 2] and u, v ~ U[-1, 1]; t_end 0.25."""
    rng = rng_for("riemann", index)
    xs, ys = rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7)
    states = np.array([[_log_uniform(rng, 0.1, 2.0), rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0),
                        _log_uniform(rng, 0.1, 2.0)] for _ in range(4)])
    X, Y = np.meshgrid(_ref_centres(n), _ref_centres(n))
    q = (X > xs).astype(int) + 2 * (Y > ys).astype(int)
    U = to_conserved(*(states[q, k] for k in range(4)))
    return U, 0.0, 0.25, {"split_x": xs, "split_y": ys, "states": states}


def ref_blast(index: int, n: int):
    """data/families.py:82-119: background rho0, p0 (log-uniform in [0.5, 2], [0.1, 0.5]); 1-3 blasts
    (centre ~ U[0, 1]^2, radius ~ U[0.04, 0.12], pressure and density ratios log-uniform in [5, 40] and
    [0.5, 2], added to the background); 0-2 density bubbles in pressure balance (radius ~ U[0.05, 0.15],
    ratio log-uniform in [0.3, 3], multiplying the density); tanh edges of width 0.005; t_end 0.25."""
    rng = rng_for("blast", index)
    rho0, p0 = _log_uniform(rng, 0.5, 2.0), _log_uniform(rng, 0.1, 0.5)
    X, Y = np.meshgrid(_ref_centres(n), _ref_centres(n))
    rho, p = np.full((n, n), rho0), np.full((n, n), p0)
    blasts, bubbles = [], []
    for _ in range(rng.integers(1, 4)):
        cx, cy = rng.uniform(0, 1, size=2)
        radius = rng.uniform(0.04, 0.12)
        pr, rr = _log_uniform(rng, 5.0, 40.0), _log_uniform(rng, 0.5, 2.0)
        w = _ref_disc(X, Y, cx, cy, radius)
        p += w * (pr - 1.0) * p0
        rho += w * (rr - 1.0) * rho0
        blasts.append([cx, cy, radius, pr, rr])
    for _ in range(rng.integers(0, 3)):
        cx, cy = rng.uniform(0, 1, size=2)
        radius = rng.uniform(0.05, 0.15)
        br = _log_uniform(rng, 0.3, 3.0)
        rho *= 1.0 + _ref_disc(X, Y, cx, cy, radius) * (br - 1.0)
        bubbles.append([cx, cy, radius, br])
    z = np.zeros((n, n))
    return to_conserved(rho, z, z, p), 0.0, 0.25, {"rho0": rho0, "p0": p0, "blasts": blasts, "bubbles": bubbles}


def ref_kh(index: int, n: int):
    """data/families.py:137-171: shear strip of a random phase y0, outside density log-uniform in [0.5, 2],
    strip/outside ratio log-uniform in [0.33, 3], shear ~ U[0.5, 1.5], bulk drift ~ U[-0.3, 0.3], pressure
    log-uniform in [1.5, 3.5], layer thickness ~ U[0.015, 0.04], 1-4 seed wavelengths of amplitude
    ~ U[0.005, 0.02] and random phase; recorded from t_start ~ U[0.2, 1.5] for 0.25."""
    rng = rng_for("kh", index)
    rho1 = _log_uniform(rng, 0.5, 2.0)
    rho2 = rho1 * _log_uniform(rng, 0.33, 3.0)
    shear = rng.uniform(0.5, 1.5)
    drift = rng.uniform(-0.3, 0.3)
    p0 = _log_uniform(rng, 1.5, 3.5)
    L = rng.uniform(0.015, 0.04)
    waves = int(rng.integers(1, 5))
    amp = rng.uniform(0.005, 0.02)
    phase = rng.uniform(0, 2 * np.pi)
    y0 = rng.uniform(0, 1)
    t_start = rng.uniform(0.2, 1.5)
    X, Y = np.meshgrid(_ref_centres(n), _ref_centres(n))
    yy = (Y - y0) % 1.0
    with np.errstate(over="ignore"):  # every branch is evaluated; the one kept is finite
        s = np.where(yy < 0.25, 0.5 * np.exp((yy - 0.25) / L),
                     np.where(yy < 0.5, 1.0 - 0.5 * np.exp((0.25 - yy) / L),
                              np.where(yy < 0.75, 1.0 - 0.5 * np.exp((yy - 0.75) / L),
                                       0.5 * np.exp((0.75 - yy) / L))))
    rho = rho1 + (rho2 - rho1) * s
    u = drift + 0.5 * shear * (1.0 - 2.0 * s)
    v = amp * np.sin(2 * np.pi * waves * X + phase)
    U = to_conserved(rho, u, v, np.full((n, n), p0))
    return U, t_start, t_start + 0.25, {"waves": waves, "t_start": t_start, "y0": y0}


REFERENCE_FAMILIES = {"riemann": ref_riemann, "blast": ref_blast, "kh": ref_kh}

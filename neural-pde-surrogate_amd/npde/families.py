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

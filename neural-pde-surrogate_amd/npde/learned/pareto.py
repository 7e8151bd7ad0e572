"""Accuracy versus cost of the plain GPU solver ladder and the learned coarse solver: the GPU counterpart of
the reference's experiments/solver_baseline.py (the plain solver as a baseline, scored on the 64^2 grid
against the block-averaged 512^2 truth, :96-124) and experiments/pareto.py (every method as one point of
accuracy against cost per 50-frame trajectory, :45-56).

  python -m npde.learned.pareto --out DIR [--instances 64] [--ladder 32,64,128,256] [--no-learned]

Instances: held-out random-Riemann instances (indices 100000 + k, never trained on), t = 0 to 0.5, 50 frames.
For every method, on the same instances:
  cost          seconds per trajectory: the wall time of the batch's device-resident run (all instances at
                once, graph capture outside the timing) divided by the instance count; the learned solver's
                includes its failure decision and any fallback fine solves
  rmse_last     the reference's Pareto accuracy: median over instances of the last frame's RMSE over the
                truth's per-variable standard deviation, averaged over the four variables
  l1_rho        mean over frames 1..50 and instances of the density L1 (the learned part's metric)
  drift_max     largest |drift| of any total over the frames (surrogate/rollout.py:54-62)
  bad_cells     cells with rho <= 0 or p not > 0 over all frames
Resolutions below 64 are scored piecewise constant on the 64^2 grid, as the reference does for 32^2.
Writes DIR/pareto.json.
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np
import torch

from ..surrogates import rollout_metrics
from . import pipeline as P


def _metrics(pred: np.ndarray, truth: np.ndarray, std: np.ndarray) -> dict:
    """pred, truth (N, 51, 4, 64, 64) float32 -> the table's accuracy columns (frames 1..50)."""
    last, l1, drift, bad = [], [], 0.0, 0
    for p, t in zip(pred, truth):
        m = rollout_metrics(torch.from_numpy(p[1:].astype(np.float64)), torch.from_numpy(t[1:].astype(np.float64)),
                            std)
        last.append(float(m["rmse"][-1].mean()))
        l1.append(float(m["l1_rho"].mean()))
        drift = max(drift, float(m["drift"].abs().max()))
        bad += int(m["bad_cells"].sum())
    return {"rmse_last": float(np.median(last)), "l1_rho": float(np.mean(l1)), "drift_max": drift, "bad_cells": bad}


def pareto_flags(rows) -> list:
    """Per row, True when no other row is at least as cheap and at least as accurate (last-frame RMSE) and
    strictly better in one of the two: the non-dominated points of the accuracy-versus-cost plane that
    experiments/pareto.py:39-56 plots."""
    return [not any(o is not r and o["cost"] <= r["cost"] and o["rmse_last"] <= r["rmse_last"]
                    and (o["cost"] < r["cost"] or o["rmse_last"] < r["rmse_last"]) for o in rows) for r in rows]


def run(instances=64, ladder=(32, 64, 128, 256), learned=True, log=print) -> dict:
    idx = list(range(100000, 100000 + instances))
    t0 = time.perf_counter()
    truth, _ = P.generate(idx, sizes=(P.COARSE,), log=lambda *a: None)
    truth = truth[P.COARSE]
    log(f"truth: {instances} instances at {P.FINE}^2 block-averaged to {P.COARSE}^2 "
        f"({time.perf_counter() - t0:.1f} s)")
    std = truth[:, 1:].astype(np.float64).transpose(2, 0, 1, 3, 4).reshape(4, -1).std(axis=1)
    rows = []
    for n in ladder:
        m = min(n, P.COARSE)
        frames, secs = P.plain_rollout(idx, n, m=m)
        if m < P.COARSE:
            f = P.COARSE // m
            frames = np.repeat(np.repeat(frames, f, axis=-2), f, axis=-1)
        row = {"method": f"plain {n}^2", "n": n, "cost": secs / instances, **_metrics(frames, truth, std)}
        rows.append(row)
        log(f"  {row['method']:>14}: {row['cost'] * 1e3:8.3f} ms per trajectory  rmse_last {row['rmse_last']:.4f}  "
            f"L1(rho) {row['l1_rho']:.5f}")
    if learned:
        from .model import load_trained
        timing = {}
        frames, secs, fell = P.learned_rollout(load_trained(), idx, log=log, timing=timing)
        row = {"method": f"learned {P.COARSE}^2", "n": P.COARSE, "cost": secs / instances, "fallbacks": len(fell),
               **_metrics(frames, truth, std)}
        rows.append(row)
        log(f"  {row['method']:>14}: {row['cost'] * 1e3:8.3f} ms per trajectory  rmse_last {row['rmse_last']:.4f}  "
            f"L1(rho) {row['l1_rho']:.5f}  fallbacks {len(fell)}")
    # the Pareto front: methods no other method beats in both cost and last-frame accuracy
    for r, front in zip(rows, pareto_flags(rows)):
        r["pareto"] = front
    return {"instances": instances, "t_end": P.T_END, "frames": P.NOUT, "truth": f"{P.FINE}^2 block-averaged",
            "std": std.tolist(), "rows": rows}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--ladder", default="32,64,128,256")
    ap.add_argument("--no-learned", action="store_true")
    args = ap.parse_args(argv)
    torch.cuda.set_device(0)
    res = run(args.instances, tuple(int(s) for s in args.ladder.split(",")), not args.no_learned)
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    (out / "pareto.json").write_text(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()

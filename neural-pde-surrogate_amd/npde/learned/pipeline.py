"""Data generation, training and evaluation of the learned flux correction (random-Riemann family).

  python -m npde.learned.pipeline --out DIR [--train-traj 448] [--iters 600] ...

1. Truth: the random-Riemann generator (npde.families.riemann, indices 0..N-1) solved at 512^2 on the
   GPU in batches of 64 (HLLC + van Leer, CFL 0.5, periodic, t_end 0.5), frames every 0.01 block-
   averaged to 64^2 (float32), as the reference's dataset pipeline does with separate processes
   (data/generate.py:60-105; trajectory.hpp:31-51).
2. Training: batch 64 of (trajectory, start frame); the coarse 64^2 solver with the network's flux
   increments is unrolled over up to 8 frame intervals (each interval: one network evaluation, then
   CFL substeps landing on the frame time; the unroll length grows 2 -> 4 -> 8 over the run), loss =
   mean over intervals of the normalised L1 error of the conserved variables (density weighted double);
   every sample mapped by a random symmetry of the periodic square; AdamW with warm-up + cosine decay
   and global-norm clipping 1.0 (the reference's optimiser chain, surrogate/train.py:209-211).
3. Evaluation: held-out indices; density L1 against the block-averaged 512^2 truth at every frame,
   averaged over frames (the protocol of experiments/solver_baseline.py:96-124), for the plain GPU
   solver at 64, 128 and 256 and the learned solver at 64; wall times of each batch run; the plain
   resolution matching the learned error and the matched-accuracy speed-up, by log-log interpolation.

The learned rollout is one device-resident recorded run: the library evaluates the network (its fused
inference kernel) at every frame landing of every instance and the stage kernels add the scaled
increments, so a rollout costs one npde_run_trajectory call like the plain solver's.

Fallback: a learned instance that met a non-physical state (the solver's own positivity check, which
also catches non-finite states), or whose density rose above 40x its initial maximum at a frame landing
(the error indicator, kept on the device by the network kernel; BLOWUP below), is re-solved by the plain binary64 fine solve, whose block averages replace
it; the fallback instances of a batch are solved together in one batch.
"""
from __future__ import annotations

import argparse
import json
import math
import time
from pathlib import Path

import numpy as np
import torch
import torch.utils.checkpoint

from .. import solver as S
from .._lib import NonPhysicalStateError
from ..families import riemann
from . import torch_solver as ts
from .model import FluxCorrector

FINE, COARSE = 512, 64
T_END, NOUT = 0.5, 50
# error indicator: density above BLOWUP x its initial maximum counts as a failure.  Converging quadrant
# flows compress far beyond a single shock's 6x: over the 512 training and held-out instances of the
# random-Riemann family the plain 64^2 solver reaches up to 18x (the 512^2 truth of the same instance
# 17x), so the indicator sits at about twice the family's largest physical compression.
BLOWUP = 40.0


def frame_times(nout=NOUT, t_end=T_END):
    return [0.0 + (t_end - 0.0) * k / nout for k in range(nout + 1)]


def block_average_f64(U: np.ndarray, m: int) -> np.ndarray:
    """(4, n, n) -> (4, m, m) binary64 block means (row sums then column sums)."""
    f = U.shape[-1] // m
    return U.reshape(4, m, f, m, f).sum(axis=(2, 4)) / (f * f)


def generate(indices, n_fine=FINE, sizes=(COARSE,), batch=64, log=print):
    """Solve the generator's instances at n_fine; returns {m: float32 (N, NOUT+1, 4, m, m)} and seconds."""
    out = {m: [] for m in sizes}
    t0 = time.perf_counter()
    grids = None
    for s0 in range(0, len(indices), batch):
        idx = list(indices[s0:s0 + batch])
        if grids is None or len(grids) != len(idx):
            grids = [S.make_grid(n_fine, n_fine, bc_x=S.PERIODIC, bc_y=S.PERIODIC)] * len(idx)
        with S.Batch(grids, "hllc", "vanleer", 0.5) as b:
            for k, i in enumerate(idx):
                b.set_state(k, riemann(i, n_fine)[0])
            b.run_trajectory(T_END, NOUT, sizes=sizes)
            for m in sizes:
                out[m].append(np.stack([b.trajectory(k, m) for k in range(len(idx))]))
        log(f"  generated {s0 + len(idx)}/{len(indices)} trajectories at {n_fine}^2 "
            f"({time.perf_counter() - t0:.1f} s)")
    return {m: np.concatenate(v) for m, v in out.items()}, time.perf_counter() - t0


# per-variable divisors of the training loss's L1 errors: density counts double (the metric is density L1)
SCALES = torch.tensor([0.5, 1.0, 1.0, 2.5], dtype=torch.float64)


def dihedral(U: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
    """U (B, ..., 4, n, n) conserved states on the periodic unit square, each sample mapped by one of the 8
    symmetries of the square: g (B,) in 0..7, bit 0 mirrors x (x -> -x: rho*u changes sign), bit 1
    mirrors y (rho*v changes sign), bit 2 exchanges x and y (and rho*u with rho*v).  The Euler equations
    commute with these maps, and so does block averaging, so a solved trajectory and its block-averaged
    truth map to another solved trajectory and its truth."""
    out = U.clone()
    for b in range(U.shape[0]):
        x = out[b]
        k = int(g[b])
        if k & 1:
            x = x.flip(-1)
            x[..., 1, :, :] = -x[..., 1, :, :]
        if k & 2:
            x = x.flip(-2)
            x[..., 2, :, :] = -x[..., 2, :, :]
        if k & 4:
            x = x.transpose(-1, -2)[..., [0, 2, 1, 3], :, :]
        out[b] = x
    return out


def train(data: torch.Tensor, iters=600, batch=32, unroll=8, lr=1e-3, warmup=50, seed=0, log=print,
          model: FluxCorrector | None = None, recompute=False, curriculum=None, augment=True, max_seconds=None):
    """data: (N, NOUT+1, 4, 64, 64) float32 on the GPU.  curriculum: unroll length as a function of the
    fraction of the run done (iterations, or the time budget max_seconds if that runs out first; the
    learning-rate schedule follows the same fraction).  recompute: recompute each interval's forward in the
    backward (activation checkpointing) instead of keeping it (a batch of 64 unrolled over 8 intervals keeps
    ~100 GB).  augment: every sample of a batch is mapped by a random symmetry of the square (dihedral)."""
    dev = data.device
    model = model or FluxCorrector(seed=seed)
    model = model.to(dev)
    opt = torch.optim.AdamW(model.parameters(), lr=lr, weight_decay=1e-4)
    t_start = time.perf_counter()

    def progress(i):
        """fraction of the run done: by iterations, or by the time budget if that is spent sooner"""
        p = i / iters
        if max_seconds:
            p = max(p, (time.perf_counter() - t_start) / max_seconds)
        return min(1.0, p)

    sched = torch.optim.lr_scheduler.LambdaLR(
        opt, lambda i: min(1.0, (i + 1) / warmup) * 0.5 * (1 + math.cos(math.pi * progress(i))))
    gen = torch.Generator().manual_seed(seed)
    N, F = data.shape[0], data.shape[1]
    n = data.shape[-1]
    dx = 1.0 / n
    S = model.calls_per_frame
    # the network's call times: S per frame interval, as the device loop lands on them (frame_times of
    # S x the frame count: the frame times themselves are bitwise among them)
    tfs = torch.tensor(frame_times((F - 1) * S), dtype=torch.float64, device=dev)
    scales = SCALES.to(dev).view(1, 4, 1, 1)
    hist, ratios, skipped = [], [], 0
    t0 = time.perf_counter()
    def rollout_loss(traj, k0, nu, corrected=True, sym=None):
        """Mean over the nu intervals of the normalised L1 error; also the samples that went non-finite."""
        frames = torch.stack([data[traj, k0 + j] for j in range(nu + 1)], dim=1).to(torch.float64)
        if sym is not None:
            frames = dihedral(frames, sym)
        U = frames[:, 0]
        kd = k0.to(dev)
        t = tfs[kd * S]
        loss = 0.0
        for j in range(1, nu + 1):
            def interval(U_, t_, j_=j):
                for s_ in range(1, S + 1):
                    tn_ = tfs[(kd + j_ - 1) * S + s_]
                    if corrected:
                        cx, cy = model(U_)
                        U_, t_, _ = ts.advance(U_, t_, tn_, dx, dx, cx_corr=cx, cy_corr=cy)
                    else:
                        U_, t_, _ = ts.advance(U_, t_, tn_, dx, dx)
                return U_, t_

            if recompute and corrected:
                U, t = torch.utils.checkpoint.checkpoint(interval, U, t, use_reentrant=False)
            else:
                U, t = interval(U, t)
            target = frames[:, j]
            err = ((U - target).abs() / scales).mean(dim=(1, 2, 3))
            finite = torch.isfinite(err)
            if not bool(finite.all()):
                return None, torch.nonzero(~finite).flatten()
            loss = loss + err.mean()
        return loss / nu, None

    for it in range(iters):
        if max_seconds is not None and time.perf_counter() - t0 > max_seconds:
            log(f"  stopping after {it} iterations: the {max_seconds:.0f} s budget is spent")
            break
        nu = unroll if curriculum is None else curriculum(progress(it))
        traj = torch.randint(0, N, (batch,), generator=gen)
        k0 = torch.randint(0, F - 1 - nu + 1, (batch,), generator=gen)
        sym = torch.randint(0, 8, (batch,), generator=gen) if augment else None
        opt.zero_grad(set_to_none=True)
        loss, badk = rollout_loss(traj, k0, nu, sym=sym)
        if badk is not None:
            # a corrected rollout left the physical states: its NaNs would poison every gradient, so
            # the batch is recomputed without those samples
            log(f"  iter {it:4d}  interval state non-finite for trajectories "
                f"{[int(traj[b]) for b in badk.tolist()][:6]} from frames {[int(k0[b]) for b in badk.tolist()][:6]}")
            keep = torch.ones(batch, dtype=torch.bool)
            keep[badk.cpu()] = False
            traj, k0 = traj[keep], k0[keep]
            sym = sym[keep] if sym is not None else None
            skipped += 1
            loss, badk = (None, None) if len(traj) == 0 else rollout_loss(traj, k0, nu, sym=sym)
        if loss is None or not bool(torch.isfinite(loss)):
            sched.step()
            continue
        loss.backward()
        gnorm = torch.nn.utils.clip_grad_norm_(model.parameters(), 1.0)
        if torch.isfinite(gnorm):
            opt.step()
        sched.step()
        hist.append(float(loss.detach()))
        if it % 25 == 0 or it == iters - 1:
            with torch.no_grad():
                plain, _ = rollout_loss(traj, k0, nu, corrected=False, sym=sym)
            ratio = float(loss.detach()) / float(plain) if plain is not None else float("nan")
            log(f"  iter {it:4d}  unroll {nu}  loss {float(loss.detach()):.5f}  ratio to plain {ratio:.4f}  "
                f"|g| {float(gnorm):.3g}  lr {sched.get_last_lr()[0]:.2e}  skipped {skipped}  "
                f"({time.perf_counter() - t0:.0f} s)")
            ratios.append(ratio)
    # held-out check of the same loss: fixed (trajectory, frame) pairs, full unroll, no gradients
    vg = torch.Generator().manual_seed(12345)
    vt = torch.randint(0, N, (batch,), generator=vg)
    vk = torch.randint(0, F - 1 - unroll + 1, (batch,), generator=vg)
    with torch.no_grad():
        lv, _ = rollout_loss(vt, vk, unroll)
        pv, _ = rollout_loss(vt, vk, unroll, corrected=False)
    model.val_ratio = float(lv) / float(pv) if lv is not None and pv is not None else float("nan")
    log(f"  final loss ratio to the plain coarse solver (unroll {unroll}, {batch} fixed samples): "
        f"{model.val_ratio:.4f}")
    model.train_ratios = ratios
    return model, hist


def l1_density(pred: np.ndarray, truth: np.ndarray) -> np.ndarray:
    """(N, F, 4, m, m) pair -> (N, F) cell-averaged |rho_pred - rho_true| (area-weighted L1 on [0,1]^2)."""
    return np.abs(pred[:, :, 0].astype(np.float64) - truth[:, :, 0].astype(np.float64)).mean(axis=(-1, -2))


def _timed_run(b, ics, t_end, nout, sizes, net=False):
    """A recorded run of the batch from the states ics, after a short untimed run with the same frame
    count and sizes (graph capture and kernel loads stay outside the timing).  Returns the device-loop
    seconds; an instance that meets a non-physical state stops there (the others finish), which the
    caller reads from the instances' times.  net: the batch runs the flux network, whose indicator is
    cleared before the timed run."""
    for k, U in enumerate(ics):
        b.set_state(k, U)
    try:
        b.run_trajectory(t_end / 10, nout, sizes=sizes)
    except NonPhysicalStateError:
        pass
    for k, U in enumerate(ics):
        b.set_state(k, U)
    if net:
        b.flux_network_peak(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    try:
        b.run_trajectory(t_end, nout, sizes=sizes)
    except NonPhysicalStateError:
        pass
    return time.perf_counter() - t0


def plain_rollout(indices, n, m=COARSE, batch=64, n_fine=FINE):
    """Plain GPU solver at n from the block-averaged n_fine IC; frames block-averaged to m. Returns frames
    (N, F, 4, m, m) float32 and the device-loop seconds (sum over batches)."""
    frames, secs = [], 0.0
    for s0 in range(0, len(indices), batch):
        idx = list(indices[s0:s0 + batch])
        grids = [S.make_grid(n, n, bc_x=S.PERIODIC, bc_y=S.PERIODIC)] * len(idx)
        ics = [block_average_f64(riemann(i, n_fine)[0], n) for i in idx]
        with S.Batch(grids, "hllc", "vanleer", 0.5) as b:
            secs += _timed_run(b, ics, T_END, NOUT, (m,))
            frames.append(np.stack([b.trajectory(k, m) for k in range(len(idx))]))
    return np.concatenate(frames), secs


def learned_rollout_states(model: FluxCorrector, U0: np.ndarray, tf, force_fallback=False, info=None):
    """The learned coarse solver from the periodic states U0 (B, 4, n, n) binary64 over the frame times tf
    (tf[0] = 0: the states' own time): one device-resident recorded run with the network in the loop --
    at every frame landing (all instances land together) the library's fused network kernel turns their
    states into the face factors of the next interval, and the stage kernels scale and add them
    (npde_set_flux_network).  Returns frames (B, F, 4, n, n) float32, the seconds of the run with its
    failure decision, and the mask of instances that failed: the solver's positivity check stopped them
    (their time is short of the end), or the network saw a density above BLOWUP x the instance's initial
    maximum (the indicator the fused kernel keeps on the device).  A dict passed as `info` receives under
    "switched_off" the mask of instances whose network the library switched off during the run (their
    signal speed ran away: they finished as the plain coarse solver; npde_set_flux_network_limit)."""
    B, n = U0.shape[0], U0.shape[-1]
    nout, t_end = len(tf) - 1, float(tf[-1])
    calls = model.calls_per_frame  # the run lands calls x nout times; every calls-th landing is a frame
    rho_max = U0[:, 0].reshape(B, -1).max(axis=1)
    grids = [S.make_grid(n, n, bc_x=S.PERIODIC, bc_y=S.PERIODIC)] * B
    with S.Batch(grids, "hllc", "vanleer", 0.5) as b:
        b.set_flux_network(model)
        secs = _timed_run(b, list(U0), t_end, nout * calls, (n,), net=True)
        t0 = time.perf_counter()
        t_now, _ = b.times()
        peak = b.flux_network_peak()
        failed = (t_now != t_end) | (peak > BLOWUP * rho_max)
        if force_fallback:
            failed[:] = True
        secs += time.perf_counter() - t0
        frames = np.stack([b.trajectory(k, n)[::calls] for k in range(B)])
        if info is not None:
            info["switched_off"] = b.flux_network_switched_off()
    return frames, secs, failed


def learned_rollout(model: FluxCorrector, indices, n=COARSE, batch=64, force_fallback=False, log=print,
                    timing=None, n_fine=FINE):
    """Random-Riemann instances: the learned coarse solver at n from the block-averaged n_fine initial states;
    frames (N, F, 4, n, n) float32, seconds, and the indices that fell back to the plain fine solve
    (whose frames replace theirs; its time is counted).  The fallback instances of a batch are solved
    together, in one batch of the plain solver.  A dict passed as `timing` receives the seconds of the
    fallback solves alone under "fallback", and the indices whose network the library switched off (they
    keep their own frames) under "switched_off"."""
    tf = frame_times()
    frames_all, secs, fell, switched = [], 0.0, [], []
    fallback_secs = 0.0
    for s0 in range(0, len(indices), batch):
        idx = list(indices[s0:s0 + batch])
        U0 = np.stack([block_average_f64(riemann(i, n_fine)[0], n) for i in idx])
        info = {}
        out, sec, badn = learned_rollout_states(model, U0, tf, force_fallback, info=info)
        switched += [idx[k] for k in np.nonzero(info["switched_off"] & ~badn)[0]]
        secs += sec
        if badn.any():
            redo = [idx[k] for k in np.nonzero(badn)[0]]
            t1 = time.perf_counter()
            fine, _ = generate(redo, n_fine=n_fine, sizes=(n,), log=lambda *a: None)
            secs += time.perf_counter() - t1
            fallback_secs += time.perf_counter() - t1
            for r, k in enumerate(np.nonzero(badn)[0]):
                out[k] = fine[n][r]
            fell += redo
            log(f"  fallback to the fine solve for {len(redo)} instance(s): {redo[:8]} "
                f"({time.perf_counter() - t1:.3f} s of the total)")
        frames_all.append(out)
    if timing is not None:
        timing["fallback"] = fallback_secs
        timing["switched_off"] = switched
    if switched:
        log(f"  network switched off (runaway signal speed; plain coarse solver from then on) for "
            f"{len(switched)} instance(s): {switched[:8]}")
    return np.concatenate(frames_all), secs, fell


def matched_speedup(ns, errs, times, e_target, t_learned):
    """Plain resolution whose error equals e_target (log-log interpolation in n) and its time."""
    ln, le, lt = np.log(ns), np.log(errs), np.log(times)
    if e_target >= errs[0]:
        return float(ns[0]), float(times[0]), float(times[0] / t_learned)
    for a in range(len(ns) - 1):
        if errs[a] >= e_target >= errs[a + 1]:
            w = (math.log(e_target) - le[a]) / (le[a + 1] - le[a])
            n_star = math.exp(ln[a] + w * (ln[a + 1] - ln[a]))
            t_star = math.exp(lt[a] + w * (lt[a + 1] - lt[a]))
            return n_star, t_star, t_star / t_learned
    # beyond the finest plain level: extrapolate with the last segment's slopes
    a = len(ns) - 2
    w = (math.log(e_target) - le[a]) / (le[a + 1] - le[a])
    n_star = math.exp(ln[a] + w * (ln[a + 1] - ln[a]))
    t_star = math.exp(lt[a] + w * (lt[a + 1] - lt[a]))
    return n_star, t_star, t_star / t_learned


def evaluate(model, test_idx, truth64, ladder=(64, 128, 256), log=print, n=COARSE, n_fine=FINE, batch=64):
    """The matched-accuracy comparison on held-out instances: the plain solver at every ladder resolution and
    the learned solver at n, all scored at n against truth64 (the n_fine truth block-averaged to n, frames
    (N, NOUT+1, 4, n, n)), every solver run on batches of `batch` instances; the learned row is stored under
    "learned64" for n = 64, else "learned<n>"."""
    res = {"test_instances": len(test_idx), "n": n, "n_fine": n_fine, "batch": batch, "plain": {}}
    for k in ladder:
        fr, secs = plain_rollout(test_idx, k, m=min(k, n), n_fine=n_fine, batch=batch)
        if k < n:  # scored piecewise constant on the n grid
            f = n // k
            fr = np.repeat(np.repeat(fr, f, axis=-2), f, axis=-1)
        if k == n:
            plain_n = fr
        e = float(l1_density(fr, truth64)[:, 1:].mean())
        res["plain"][k] = {"l1_rho": e, "seconds": secs}
        log(f"  plain {k:4d}^2: L1(rho) {e:.5f}  wall {secs:.3f} s")
    timing = {}
    fr, secs, fell = learned_rollout(model, test_idx, n=n, log=log, timing=timing, n_fine=n_fine, batch=batch)
    e = float(l1_density(fr, truth64)[:, 1:].mean())
    own = np.array([i not in fell for i in test_idx])
    e_own = float(l1_density(fr[own], truth64[own])[:, 1:].mean())
    e_plain_own = float(l1_density(plain_n[own], truth64[own])[:, 1:].mean())
    res[f"learned{n}"] = {"l1_rho": e, "seconds": secs, "fallbacks": len(fell), "fallback_seconds": timing["fallback"],
                          "network_switched_off": len(timing["switched_off"]),
                          "l1_rho_without_fallback_instances": e_own,
                          f"plain{n}_l1_rho_same_instances": e_plain_own}
    log(f"  learned {n}^2: L1(rho) {e:.5f}  wall {secs:.3f} s  fallbacks {len(fell)}  "
        f"(instances that kept the learned solver: {e_own:.5f} against plain {n}^2 {e_plain_own:.5f})")
    ns = np.array(ladder, dtype=float)
    errs = np.array([res["plain"][n]["l1_rho"] for n in ladder])
    times = np.array([res["plain"][n]["seconds"] for n in ladder])
    n_star, t_star, sp = matched_speedup(ns, errs, times, e, secs)
    res["matched"] = {"plain_resolution": n_star, "plain_seconds": t_star, "speedup": sp}
    log(f"  matched plain resolution {n_star:.1f}^2 ({t_star:.3f} s): learned is {sp:.2f}x faster")
    return res


def evaluate_kh(model, indices, n=128, n_fine=1024, log=print, force_fallback=False, keep_frames=False):
    """Held-out Kelvin-Helmholtz instances (npde.families.kh, never seen in training): from a random start
    time t_s in [0.5, 1.5] to t = 2.  Truth and fallback: the plain binary64 1024^2 solve from t = 0,
    block-averaged to n at NOUT frames after t_s (single-instance runs; their [t_s, 2] wall time is the
    fine solver's cost; times of all runs count from t_s).  The plain and learned solvers at n start from the truth's state at t_s,
    block-averaged in binary64, and run the same NOUT frame intervals; an instance whose learned run
    fails takes the fine frames and is charged the fine solve's time.  force_fallback marks every learned
    run as failed (the fallback path itself); keep_frames adds each instance's truth, plain and learned
    frames (NOUT+1, 4, n, n) float32 to its row."""
    from ..families import kh
    rows = []
    for i in indices:
        U, t_end, meta = kh(i, n_fine)
        ts = float(meta["t_start"])
        with S.Batch([S.make_grid(n_fine, n_fine, bc_x=S.PERIODIC, bc_y=S.PERIODIC)], "hllc", "vanleer", 0.5) as b:
            b.set_state(0, U)
            b.advance_to(ts)
            fine_ts = b.get_state(0)
            # from here on times count from t_s (set_state restarts the clock), as in the coarse runs
            b.set_state(0, fine_ts)
            b.run_trajectory(0.001, 1, sizes=(n,))  # warm-up: graph capture outside the timing
            b.set_state(0, fine_ts)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.run_trajectory(t_end - ts, NOUT, sizes=(n,))
            fine_s = time.perf_counter() - t0
            truth = b.trajectory(0, n)[None]
        U0 = block_average_f64(fine_ts, n)[None]
        tf = frame_times(t_end=t_end - ts)
        with S.Batch([S.make_grid(n, n, bc_x=S.PERIODIC, bc_y=S.PERIODIC)], "hllc", "vanleer", 0.5) as b:
            b.set_state(0, U0[0])
            b.run_trajectory(0.01, 1, sizes=(n,))
            b.set_state(0, U0[0])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.run_trajectory(t_end - ts, NOUT, sizes=(n,))
            plain_s = time.perf_counter() - t0
            plain = b.trajectory(0, n)[None]
        lf, learned_s, bad = learned_rollout_states(model, U0, tf, force_fallback=force_fallback)
        fell = bool(bad[0])
        if fell:
            lf, learned_s = truth, learned_s + fine_s
        r = {"index": i, "t_start": ts, "fine_seconds": fine_s, "plain_seconds": plain_s,
             "learned_seconds": learned_s, "fallback": fell,
             "l1_rho_plain": float(l1_density(plain, truth)[:, 1:].mean()),
             "l1_rho_learned": float(l1_density(lf, truth)[:, 1:].mean())}
        if keep_frames:
            r["frames"] = {"truth": truth[0], "plain": plain[0], "learned": lf[0]}
        rows.append(r)
        log(f"  kh {i}: t_s {ts:.3f}  L1(rho) plain {n}^2 {r['l1_rho_plain']:.5f}  learned {r['l1_rho_learned']:.5f}"
            f"{' (fallback)' if fell else ''}  wall fine {fine_s:.3f} s, plain {plain_s:.3f} s, learned {learned_s:.3f} s")
    mean = lambda k: float(np.mean([r[k] for r in rows]))
    res = {"instances": rows, "n": n, "n_fine": n_fine,
           "mean": {k: mean(k) for k in ("l1_rho_plain", "l1_rho_learned", "fine_seconds", "plain_seconds",
                                        "learned_seconds")},
           "fallbacks": int(sum(r["fallback"] for r in rows))}
    m = res["mean"]
    log(f"  kh mean: L1(rho) plain {m['l1_rho_plain']:.5f}, learned {m['l1_rho_learned']:.5f}; wall fine "
        f"{m['fine_seconds']:.3f} s, plain {m['plain_seconds']:.3f} s, learned {m['learned_seconds']:.3f} s; "
        f"fallbacks {res['fallbacks']}")
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--train-traj", type=int, default=448)
    ap.add_argument("--test-traj", type=int, default=64)
    ap.add_argument("--iters", type=int, default=600)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--unroll", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init", default=None, help="start from these weights (.npz) at the full unroll")
    ap.add_argument("--max-seconds", type=float, default=None, help="stop training after this many seconds")
    ap.add_argument("--out-scale", type=float, default=None, help="bound of the increments (fraction of the face "
                    "flux scale) for a new corrector")
    ap.add_argument("--calls-per-frame", type=int, default=None, help="network evaluations per frame interval "
                    "(a new corrector, or a change for the one given by --init)")
    ap.add_argument("--kh-eval", type=int, default=0, help="only evaluate the trained weights on this many held-out "
                    "Kelvin-Helmholtz instances at 128^2 (kh128_eval.json)")
    ap.add_argument("--rollout", choices=("plain", "learned"), default=None, help="only time the 64^2 rollout of "
                    "the held-out instances (--test-traj of them, in one batch) three times, plain or learned "
                    "(rollout_times.json): a workload for a kernel trace of either loop alone")
    args = ap.parse_args(argv)
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    logf = open(out / "learned_log.txt", "w")

    def log(*a):
        s = " ".join(str(x) for x in a)
        print(s, flush=True)
        logf.write(s + "\n")
        logf.flush()

    torch.cuda.set_device(0)
    if args.rollout:
        idx = list(range(100000, 100000 + args.test_traj))
        U0 = np.stack([block_average_f64(riemann(i, FINE)[0], COARSE) for i in idx])
        grids = [S.make_grid(COARSE, COARSE, bc_x=S.PERIODIC, bc_y=S.PERIODIC)] * len(idx)
        rows = []
        with S.Batch(grids, "hllc", "vanleer", 0.5) as b:
            if args.rollout == "learned":
                from .model import load_trained
                b.set_flux_network(load_trained(args.init) if args.init else load_trained())
            for rep in range(3):
                secs = _timed_run(b, list(U0), T_END, NOUT, (COARSE,), net=args.rollout == "learned")
                steps = b.times()[1]
                rows.append({"seconds": secs, "steps_mean": float(steps.mean()), "steps_max": int(steps.max())})
                log(f"{args.rollout} {COARSE}^2 rollout of {len(idx)} instances: {secs * 1e3:.2f} ms, steps mean "
                    f"{steps.mean():.1f}, max {steps.max()}")
        (out / "rollout_times.json").write_text(json.dumps({"rollout": args.rollout, "runs": rows}, indent=1))
        return
    if args.kh_eval:
        from .model import load_trained
        log(f"held-out Kelvin-Helmholtz evaluation: {args.kh_eval} instances")
        res = evaluate_kh(load_trained(args.init) if args.init else load_trained(), list(range(args.kh_eval)), log=log)
        (out / "kh128_eval.json").write_text(json.dumps(res, indent=1))
        return
    log(f"generating {args.train_traj} training trajectories")
    train_idx = list(range(args.train_traj))
    test_idx = list(range(100000, 100000 + args.test_traj))
    data, gen_s = generate(train_idx, log=log)
    data64 = torch.from_numpy(data[COARSE]).cuda()
    log(f"training: {args.iters} iterations, batch {args.batch}, unroll {args.unroll}")
    t0 = time.perf_counter()
    cur = lambda p: max(2, min(args.unroll, 2 ** (1 + int(3 * p))))  # 2, 4, 8 over thirds of the run
    init = None
    if args.out_scale is not None or args.calls_per_frame is not None:
        init = FluxCorrector(out_scale=0.02 if args.out_scale is None else args.out_scale,
                             calls_per_frame=args.calls_per_frame or 1)
    if args.init:  # continue a previous run: full unroll from the start
        from .model import load_trained
        init, cur = load_trained(args.init), None
        if args.calls_per_frame:
            init.calls_per_frame = args.calls_per_frame
        log(f"starting from {args.init}")
    log(f"network calls per frame interval: {(init or FluxCorrector()).calls_per_frame}")
    model, hist = train(data64, iters=args.iters, batch=args.batch, unroll=args.unroll, lr=args.lr, log=log,
                        curriculum=cur, model=init, max_seconds=args.max_seconds)
    train_s = time.perf_counter() - t0
    np.savez(out / "flux_corrector_riemann64.npz", **{k: v.numpy() for k, v in model.state().items()})
    del data64
    log("evaluation: generating held-out truth")
    truth, _ = generate(test_idx, log=log)
    res = evaluate(model, test_idx, truth[COARSE], log=log)
    res.update({"train_trajectories": args.train_traj, "iters": args.iters, "batch": args.batch,
                "unroll": args.unroll, "generate_seconds": gen_s, "train_seconds": train_s,
                "loss_first": hist[0], "loss_last": float(np.mean(hist[-20:])),
                "train_loss_ratio_to_plain": model.val_ratio,
                "parameters": model.num_parameters()})
    (out / "learned_eval.json").write_text(json.dumps(res, indent=1))
    log(json.dumps(res))


if __name__ == "__main__":
    main()

"""Data generation, training and evaluation of the learned flux correction (random-Riemann family).

  python -m npde.learned.pipeline --out DIR [--train-traj 448] [--iters 600] ...

1. Truth: the random-Riemann generator (npde.families.riemann, indices 0..N-1) solved at 512^2 on the
   GPU in batches of 64 (HLLC + van Leer, CFL 0.5, periodic, t_end 0.5), frames every 0.01 block-
   averaged to 64^2 (float32), as the reference's dataset pipeline does with separate processes
   (data/generate.py:60-105; trajectory.hpp:31-51).
2. Training: batch 32 of (trajectory, start frame); the coarse 64^2 solver with the network's flux
   increments is unrolled over up to 8 frame intervals (each interval: one network evaluation, then
   CFL substeps landing on the frame time; the unroll length grows 2 -> 4 -> 8 over the run), loss =
   mean over intervals of the normalised L1 error of the conserved variables; the backward
   recomputes each interval's forward (activation checkpointing); AdamW with warm-up + cosine decay
   and global-norm clipping 1.0 (the reference's optimiser chain, surrogate/train.py:209-211).
3. Evaluation: held-out indices; density L1 against the block-averaged 512^2 truth at every frame,
   averaged over frames (the protocol of experiments/solver_baseline.py:96-124), for the plain GPU
   solver at 64, 128 and 256 and the learned solver at 64; wall times of each batch run; the plain
   resolution matching the learned error and the matched-accuracy speed-up, by log-log interpolation.

Fallback: after every interval a learned instance that met a non-physical state (the solver's own
positivity check), became non-finite, or whose density rose above 20x its initial maximum (the error
indicator: a single ideal-gas shock compresses at most 6x at gamma = 1.4) is re-solved by the plain
binary64 fine solve, whose block averages replace it.
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
from .conv import pack_stack
from .model import FluxCorrector

FINE, COARSE = 512, 64
T_END, NOUT = 0.5, 50
BLOWUP = 20.0  # error indicator: density above BLOWUP x its initial maximum counts as a failure


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


SCALES = torch.tensor([1.0, 1.0, 1.0, 2.5], dtype=torch.float64)


def train(data: torch.Tensor, iters=600, batch=32, unroll=8, lr=1e-3, warmup=50, seed=0, log=print,
          model: FluxCorrector | None = None, recompute=True, curriculum=None):
    """data: (N, NOUT+1, 4, 64, 64) float32 on the GPU."""
    dev = data.device
    model = model or FluxCorrector(seed=seed)
    model = model.to(dev)
    opt = torch.optim.AdamW(model.parameters(), lr=lr, weight_decay=1e-4)
    sched = torch.optim.lr_scheduler.LambdaLR(
        opt, lambda i: min(1.0, (i + 1) / warmup) * 0.5 * (1 + math.cos(math.pi * min(1.0, i / iters))))
    gen = torch.Generator().manual_seed(seed)
    N, F = data.shape[0], data.shape[1]
    n = data.shape[-1]
    dx = 1.0 / n
    tf = torch.tensor(frame_times(F - 1), dtype=torch.float64, device=dev)
    scales = SCALES.to(dev).view(1, 4, 1, 1)
    hist, ratios, skipped = [], [], 0
    t0 = time.perf_counter()
    def rollout_loss(traj, k0, nu, corrected=True):
        """Mean over the nu intervals of the normalised L1 error; also the samples that went non-finite."""
        U = data[traj, k0].to(torch.float64)
        t = tf[k0.to(dev)]
        loss = 0.0
        for j in range(1, nu + 1):
            tn = tf[(k0 + j).to(dev)]

            def interval(U_, t_, tn_=tn):
                if corrected:
                    cx, cy = model(U_)
                    U2, t2, _ = ts.advance(U_, t_, tn_, dx, dx, cx_corr=cx, cy_corr=cy)
                else:
                    U2, t2, _ = ts.advance(U_, t_, tn_, dx, dx)
                return U2, t2

            if recompute and corrected:
                U, t = torch.utils.checkpoint.checkpoint(interval, U, t, use_reentrant=False)
            else:
                U, t = interval(U, t)
            target = data[traj, k0 + j].to(torch.float64)
            err = ((U - target).abs() / scales).mean(dim=(1, 2, 3))
            finite = torch.isfinite(err)
            if not bool(finite.all()):
                return None, torch.nonzero(~finite).flatten()
            loss = loss + err.mean()
        return loss / nu, None

    for it in range(iters):
        nu = unroll if curriculum is None else curriculum(it)
        traj = torch.randint(0, N, (batch,), generator=gen)
        k0 = torch.randint(0, F - 1 - nu + 1, (batch,), generator=gen)
        opt.zero_grad(set_to_none=True)
        loss, badk = rollout_loss(traj, k0, nu)
        if badk is not None:
            # a corrected rollout left the physical states: its NaNs would poison every gradient, so
            # the batch is recomputed without those samples
            log(f"  iter {it:4d}  interval state non-finite for trajectories "
                f"{[int(traj[b]) for b in badk.tolist()][:6]} from frames {[int(k0[b]) for b in badk.tolist()][:6]}")
            keep = torch.ones(batch, dtype=torch.bool)
            keep[badk.cpu()] = False
            traj, k0 = traj[keep], k0[keep]
            skipped += 1
            loss, badk = (None, None) if len(traj) == 0 else rollout_loss(traj, k0, nu)
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
                plain, _ = rollout_loss(traj, k0, nu, corrected=False)
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


def plain_rollout(indices, n, m=COARSE, batch=64):
    """Plain GPU solver at n from the block-averaged fine IC; frames block-averaged to m. Returns frames
    (N, F, 4, m, m) float32 and the device-loop seconds (sum over batches)."""
    frames, secs = [], 0.0
    for s0 in range(0, len(indices), batch):
        idx = list(indices[s0:s0 + batch])
        grids = [S.make_grid(n, n, bc_x=S.PERIODIC, bc_y=S.PERIODIC)] * len(idx)
        ics = [block_average_f64(riemann(i, FINE)[0], n) for i in idx]
        with S.Batch(grids, "hllc", "vanleer", 0.5) as b:
            for k in range(len(idx)):
                b.set_state(k, ics[k])
            b.run_trajectory(0.01, 1, sizes=(m,))  # warm-up: captures the graph outside the timing
            for k in range(len(idx)):
                b.set_state(k, ics[k])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.run_trajectory(T_END, NOUT, sizes=(m,))
            secs += time.perf_counter() - t0
            frames.append(np.stack([b.trajectory(k, m) for k in range(len(idx))]))
    return np.concatenate(frames), secs


def learned_rollout_states(model: FluxCorrector, U0: np.ndarray, tf, force_fallback=False):
    """Coarse HIP solver with the network's increments from the periodic states U0 (B, 4, n, n)
    binary64, frames at the times tf (tf[0] = 0: the states' own time); returns frames
    (B, F, 4, n, n) float32, the seconds of the rollout and the mask of instances that failed.

    The network reads the solver's state in place (zero-copy view) and writes its increments into
    fixed buffers the stage kernels read.  Everything runs on the solver's stream: per frame interval,
    one replay of the captured network graph, one device loop to the frame time, the frame copy and
    the positivity / indicator check, with no host synchronisation besides the loop's own."""
    dev = torch.device("cuda")
    B, n = U0.shape[0], U0.shape[-1]
    nout = len(tf) - 1
    packed = pack_stack(model.params())
    grids = [S.make_grid(n, n, bc_x=S.PERIODIC, bc_y=S.PERIODIC)] * B
    with S.Batch(grids, "hllc", "vanleer", 0.5) as b, torch.no_grad():
        b.set_tuning(steps_per_graph=8)  # ~14 CFL substeps per frame interval at 64^2
        for k in range(B):
            b.set_state(k, U0[k])
        stream = torch.cuda.ExternalStream(b.stream)
        with torch.cuda.stream(stream):
            U = b.state_view()
            rho_max = U[:, 0].flatten(1).amax(1)
            CX = torch.zeros(B, 4, n, n, dtype=torch.float64, device=dev)
            CY = torch.zeros_like(CX)
            for k in range(B):
                b.set_flux_correction(k, CX[k].data_ptr(), CY[k].data_ptr())
            frames = torch.empty(B, nout + 1, 4, n, n, dtype=torch.float32, device=dev)
            frames[:, 0] = U.to(torch.float32)
            keep = torch.ones(B, dtype=torch.bool, device=dev)

            def check_and_correct():
                # positivity / indicator check of the state just reached, then the increments of the
                # next interval (zero for instances that failed): one captured graph per interval
                rho = U[:, 0]
                p = (1.4 - 1.0) * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / U[:, 0])
                ok = (torch.isfinite(U).flatten(1).all(1) & (rho.flatten(1).amin(1) > 0)
                      & (p.flatten(1).amin(1) > 0) & (rho.flatten(1).amax(1) <= BLOWUP * rho_max))
                keep.logical_and_(ok)
                model.increments_into(U, CX, CY, keep, packed)

            check_and_correct()  # warm-up (allocations, kernel loads); U is the initial state
            b.advance_to(0.0)  # warm-up: captures the solver's graph outside the timing
            net = torch.cuda.CUDAGraph()
            with torch.cuda.graph(net):
                check_and_correct()
            keep.fill_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            net.replay()  # increments of the first interval
            for j in range(1, nout + 1):
                try:
                    b.advance_to(tf[j])
                except NonPhysicalStateError:
                    # every instance that stopped short of the frame time met a non-physical state;
                    # its remaining frames come from the fine solve, and its state is reset to a
                    # physical one so that it does not stop the following device loops again
                    t_now, _ = b.times()
                    for k in np.nonzero(t_now != tf[j])[0]:
                        keep[int(k)] = False
                        b.set_state(int(k), U0[k])
                frames[:, j] = U.to(torch.float32)
                if force_fallback:
                    keep.zero_()
                net.replay()  # check frame j, increments of interval j+1
            torch.cuda.synchronize()
            secs = time.perf_counter() - t0
            return frames.cpu().numpy(), secs, (~keep).cpu().numpy()


def learned_rollout(model: FluxCorrector, indices, n=COARSE, batch=64, force_fallback=False, log=print,
                    timing=None):
    """Random-Riemann instances: the learned coarse solver from the block-averaged fine initial states;
    frames (N, F, 4, n, n) float32, seconds, and the indices that fell back to the plain fine solve
    (whose frames replace theirs; its time is counted).  A dict passed as `timing` receives the seconds
    of the fallback solves alone under "fallback"."""
    tf = frame_times()
    frames_all, secs, fell = [], 0.0, []
    fallback_secs = 0.0
    for s0 in range(0, len(indices), batch):
        idx = list(indices[s0:s0 + batch])
        U0 = np.stack([block_average_f64(riemann(i, FINE)[0], n) for i in idx])
        out, sec, badn = learned_rollout_states(model, U0, tf, force_fallback)
        secs += sec
        if badn.any():
            redo = [idx[k] for k in np.nonzero(badn)[0]]
            t1 = time.perf_counter()
            fine, _ = generate(redo, sizes=(n,), log=lambda *a: None)
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


def evaluate(model, test_idx, truth64, ladder=(64, 128, 256), log=print):
    res = {"test_instances": len(test_idx), "plain": {}}
    for n in ladder:
        fr, secs = plain_rollout(test_idx, n)
        if n == COARSE:
            plain64 = fr
        e = float(l1_density(fr, truth64)[:, 1:].mean())
        res["plain"][n] = {"l1_rho": e, "seconds": secs}
        log(f"  plain {n:4d}^2: L1(rho) {e:.5f}  wall {secs:.3f} s")
    timing = {}
    fr, secs, fell = learned_rollout(model, test_idx, log=log, timing=timing)
    e = float(l1_density(fr, truth64)[:, 1:].mean())
    own = np.array([i not in fell for i in test_idx])
    e_own = float(l1_density(fr[own], truth64[own])[:, 1:].mean())
    e_plain_own = float(l1_density(plain64[own], truth64[own])[:, 1:].mean())
    res["learned64"] = {"l1_rho": e, "seconds": secs, "fallbacks": len(fell), "fallback_seconds": timing["fallback"],
                        "l1_rho_without_fallback_instances": e_own,
                        "plain64_l1_rho_same_instances": e_plain_own}
    log(f"  learned 64^2: L1(rho) {e:.5f}  wall {secs:.3f} s  fallbacks {len(fell)}  "
        f"(instances that kept the learned solver: {e_own:.5f} against plain 64^2 {e_plain_own:.5f})")
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
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--unroll", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init", default=None, help="start from these weights (.npz) at the full unroll")
    ap.add_argument("--out-scale", type=float, default=None, help="bound of the increments (fraction of the face "
                    "flux scale) for a new corrector")
    ap.add_argument("--kh-eval", type=int, default=0, help="only evaluate the trained weights on this many held-out "
                    "Kelvin-Helmholtz instances at 128^2 (kh128_eval.json)")
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
    cur = lambda it: max(2, min(args.unroll, 2 ** (1 + (3 * it) // max(1, args.iters))))  # 2, 4, 8
    init = None if args.out_scale is None else FluxCorrector(out_scale=args.out_scale)
    if args.init:  # continue a previous run: full unroll from the start
        from .model import load_trained
        init, cur = load_trained(args.init), None
        log(f"starting from {args.init}")
    model, hist = train(data64, iters=args.iters, batch=args.batch, unroll=args.unroll, lr=args.lr, log=log,
                        curriculum=cur, model=init)
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

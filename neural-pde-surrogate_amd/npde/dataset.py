"""Dataset pipeline on the GPU: generate, split and pack in one job.

  python -m npde.dataset --out DIR [--riemann 448] [--blast 96] [--kh 0] [--sizes 128,64] [--batch 256]

Replaces the reference's three steps (data/generate.py -> data/splits.py -> data/pack.py), whose
per-instance euler2d processes and raw files become batched device loops writing the packed files
directly:

1. Generate.  Instances of the benchmark families (npde.families) are solved with the plain binary64
   solver (HLLC + van Leer, CFL 0.5) in batches of up to `batch` like instances (sorted by time window
   and grid size, so that a batch's members step for similar counts), 50 frame
   intervals from the family's t_start to t_end, block-averaged on the GPU to every requested size
   (trajectory.hpp:31-51).
2. Split by (family, index) only (data/splits.py:27-33): training families (riemann, blast) go to
   train / val / test for index % 10 in 0-7 / 8 / 9; held-out families (kh) go to ood.
3. Pack, per resolution n and split, <split>_<n>.h5 in the layout of data/pack.py:55-80 (U (N, 51, 4,
   n, n) float32 in chunks of one frame, t, family, index; written by libnpde_io.so), with the checks of
   data/pack.py:66-75 (positivity; mass and energy drift of periodic boxes) and stats_<n>.json from the
   train split only (data/pack.py:96-113): per-variable mean and std of U, of log rho and of one-frame
   increments.  manifest.json lists every trajectory with its split, solver size and step count.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import time
from pathlib import Path

import numpy as np

from . import families, solver as S
from ._lib import PKG_ROOT

GAMMA = 1.4
NOUT = 50
SPLITS = ["train", "val", "test", "ood"]
VARS = ["rho", "rho_u", "rho_v", "E"]
TRAINING_FAMILIES = {"riemann", "blast"}
HELD_OUT_FAMILIES = {"kh"}
KH_N = 1024  # the held-out KH family's plain-solver resolution


def split_for(family: str, index: int) -> str:
    """data/splits.py:27-33: depends on (family, index) only, exact 80/10/10 over any index range."""
    if family in HELD_OUT_FAMILIES:
        return "ood"
    if family not in TRAINING_FAMILIES:
        raise ValueError(f"family '{family}' is neither a training nor a held-out family")
    r = index % 10
    return "train" if r < 8 else ("val" if r == 8 else "test")


class RunningStats:
    """Per-channel mean and standard deviation accumulated in binary64 over many arrays."""

    def __init__(self, channels: int):
        self.n = 0
        self.s = np.zeros(channels)
        self.ss = np.zeros(channels)

    def add(self, x: np.ndarray, axis):
        x = x.astype(np.float64)
        self.n += x.size // x.shape[1]
        self.s += x.sum(axis=axis)
        self.ss += (x ** 2).sum(axis=axis)

    def result(self):
        mean = self.s / self.n
        return mean, np.sqrt(np.maximum(self.ss / self.n - mean ** 2, 0.0))


def pressure(U: np.ndarray) -> np.ndarray:
    rho, mx, my, E = U[:, 0], U[:, 1], U[:, 2], U[:, 3]
    return (GAMMA - 1) * (E - 0.5 * (mx ** 2 + my ** 2) / rho)


REFERENCE_N = 512  # the reference dataset's solver size (data/generate.py; manifest solver_n)


def instance_size(family: str, index: int, riemann_n: int, source: str = "benchmark") -> int:
    """Grid size of an instance, without building its initial condition."""
    if source == "reference":
        return REFERENCE_N
    if family == "riemann":
        return riemann_n
    if family == "blast":
        return families.blast_size(index)
    return KH_N


def instance(family: str, index: int, riemann_n: int, source: str = "benchmark"):
    """(U, n, t_start, t_end) of one instance: of the benchmark's families (npde.families), or with
    source "reference" of the reference's dataset families at its solver size (REFERENCE_FAMILIES)."""
    if source == "reference":
        U, ts, te, _ = families.REFERENCE_FAMILIES[family](index, REFERENCE_N)
        return U, REFERENCE_N, ts, te
    if family == "riemann":
        U, t_end, _ = families.riemann(index, riemann_n)
        return U, riemann_n, 0.0, t_end
    if family == "blast":
        U, t_end, meta = families.blast(index)
        return U, meta["n"], 0.0, t_end
    if family == "kh":
        U, t_end, meta = families.kh(index, KH_N)
        return U, KH_N, meta["t_start"], t_end
    raise ValueError(f"unknown family '{family}'")


def generate(items, sizes, riemann_n=512, batch=256, log=print, source="benchmark"):
    """items: [(family, index)].  Returns {(family, index): {"frames": {m: (51, 4, m, m) float32},
    "t": (51,), "n": n, "steps": steps}} and the seconds spent.  source: "benchmark" or "reference"
    (instance)."""
    out = {}
    t0 = time.perf_counter()
    # batches of like instances (same time window, same size) step for similar counts, so a batch does
    # not wait long for its slowest member; results are keyed by (family, index), so order is free
    items = sorted(items, key=lambda it: (it[0] == "kh", instance_size(it[0], it[1], riemann_n, source)))
    for s0 in range(0, len(items), batch):
        chunk = items[s0:s0 + batch]
        ics = [instance(f, i, riemann_n, source) for f, i in chunk]
        # one batch per (t_start, t_end): instances of a batch share the frame schedule
        groups = {}
        for k, (U, n, ts, te) in enumerate(ics):
            groups.setdefault((ts, te), []).append(k)
        for (ts, te), ks in groups.items():
            grids = [S.make_grid(ics[k][1], ics[k][1], bc_x=S.PERIODIC, bc_y=S.PERIODIC) for k in ks]
            with S.Batch(grids, "hllc", "vanleer", 0.5) as b:
                for j, k in enumerate(ks):
                    b.set_state(j, ics[k][0])
                b.run_trajectory(te, NOUT, sizes=sizes, t_start=ts)
                for j, k in enumerate(ks):
                    out[chunk[k]] = {"frames": {m: b.trajectory(j, m) for m in sizes}, "t": b.frame_times(j),
                                     "n": ics[k][1], "steps": int(b.time(j)[1])}
        log(f"  generated {s0 + len(chunk)}/{len(items)} trajectories ({time.perf_counter() - t0:.1f} s)")
    return out, time.perf_counter() - t0


IO_LIB_PATH = PKG_ROOT / "lib" / "libnpde_io.so"


def _io():
    # NPDE_IO_LIB: another build of the writer (an instrumented one, for example)
    lib = C.CDLL(os.environ.get("NPDE_IO_LIB", str(IO_LIB_PATH)))
    lib.npde_io_available.restype = C.c_int
    lib.npde_io_last_error.restype = C.c_char_p
    lib.npde_io_write_packed.restype = C.c_int
    lib.npde_io_write_packed.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                         C.POINTER(C.c_char_p), C.c_void_p, C.c_char_p, C.c_double]
    if not lib.npde_io_available():
        raise RuntimeError("libnpde_io.so was built without HDF5")
    return lib


def frame_spacing(t: np.ndarray) -> float:
    """The frame spacing every row of t (N, F) shares -- the reference stores its constant 0.005 as the
    packed files' dt (data/pack.py:80) -- or NaN when rows have windows of their own (the held-out KH
    instances each start at their own t_start), in which case the file carries no dt attribute and its
    t dataset is authoritative."""
    d = np.diff(t, axis=1)
    if d.size == 0:
        return float("nan")
    ref = float(d[0, 0])
    return ref if np.allclose(d, ref, rtol=1e-9, atol=0.0) else float("nan")


def write_packed(path: Path, U: np.ndarray, t: np.ndarray, fams, idx, split: str, dt: float):
    lib = _io()
    U = np.ascontiguousarray(U, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    N, F, _, n, _ = U.shape
    names = (C.c_char_p * N)(*[f.encode() for f in fams])
    rc = lib.npde_io_write_packed(str(path).encode(), N, F, n, U.ctypes.data, t.ctypes.data, names,
                                  idx.ctypes.data, split.encode(), float(dt))
    if rc != 0:
        raise RuntimeError(f"writing {path}: {lib.npde_io_last_error().decode()}")


def pack(trajs, sizes, out_dir: Path, log=print) -> bool:
    """Write <split>_<n>.h5, stats_<n>.json and manifest.json; True when every check passes."""
    keys = sorted(trajs, key=lambda k: (k[0], k[1]))
    manifest = [{"family": f, "index": i, "split": split_for(f, i), "solver_n": trajs[(f, i)]["n"],
                 "steps": trajs[(f, i)]["steps"], "t_start": float(trajs[(f, i)]["t"][0])} for f, i in keys]
    (out_dir / "manifest.json").write_text(json.dumps({"trajectories": manifest}, indent=1))
    ok = True
    for n in sizes:
        stats = {"U": RunningStats(4), "log_rho": RunningStats(1), "dU": RunningStats(4)}
        for split in SPLITS:
            ks = [k for k in keys if split_for(*k) == split]
            if not ks:
                continue
            U = np.stack([trajs[k]["frames"][n] for k in ks])           # (N, 51, 4, n, n)
            t = np.stack([trajs[k]["t"] for k in ks])
            dt = frame_spacing(t)
            write_packed(out_dir / f"{split}_{n}.h5", U, t, [k[0] for k in ks], [k[1] for k in ks], split, dt)
            min_rho = float(U[:, :, 0].min())
            min_p = float(min(pressure(u).min() for u in U))
            totals = U.astype(np.float64).mean(axis=(3, 4))               # (N, 51, 4)
            drift = float(max((np.abs(totals[:, :, q] - totals[:, :1, q]) / np.abs(totals[:, :1, q])).max()
                              for q in (0, 3)))
            ok &= min_rho > 0 and min_p > 0 and drift < 1e-5
            log(f"  {split}_{n}.h5: {len(ks)} trajectories  min rho {min_rho:.3g}  min p {min_p:.3g}  "
                f"max mass/energy drift {drift:.1e}")
            if split == "train":
                for u in U:
                    stats["U"].add(u, axis=(0, 2, 3))
                    stats["log_rho"].add(np.log(u[:, :1]), axis=(0, 2, 3))
                    stats["dU"].add(np.diff(u, axis=0), axis=(0, 2, 3))
        if stats["U"].n:
            mean, std = stats["U"].result()
            lmean, lstd = stats["log_rho"].result()
            dmean, dstd = stats["dU"].result()
            (out_dir / f"stats_{n}.json").write_text(json.dumps({
                "source": "train split only", "n": n, "variables": VARS, "mean": mean.tolist(), "std": std.tolist(),
                "log_rho_mean": float(lmean[0]), "log_rho_std": float(lstd[0]),
                "delta_mean": dmean.tolist(), "delta_std": dstd.tolist()}, indent=1))
    return ok


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--riemann", type=int, default=448, help="riemann instances (indices 0..N-1)")
    ap.add_argument("--riemann-n", type=int, default=512)
    ap.add_argument("--blast", type=int, default=0, help="blast instances (indices 0..N-1)")
    ap.add_argument("--kh", type=int, default=0, help="kh instances (held out, 1024^2)")
    ap.add_argument("--sizes", default="128,64")
    ap.add_argument("--families", choices=("benchmark", "reference"), default="benchmark",
                    help="the benchmark's families, or the reference's dataset families at 512^2 (riemann and "
                         "blast over t = 0 .. 0.25, kh over its own t_start + 0.25)")
    ap.add_argument("--no-pack", action="store_true", help="generate (and check) only; write no files")
    ap.add_argument("--check-steps", default=None, help="a step-count fixture (tests/golden/"
                    "reference_manifest_steps.json): compare every generated trajectory's step count")
    # larger batches overlap more of their instances' tails: 160 instances took 10.8 s in batches of 64
    # and 10.1 s in one batch of 256
    ap.add_argument("--batch", type=int, default=256)
    args = ap.parse_args(argv)
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    sizes = tuple(int(s) for s in args.sizes.split(",") if s)
    items = ([("riemann", i) for i in range(args.riemann)] + [("blast", i) for i in range(args.blast)]
             + [("kh", i) for i in range(args.kh)])
    print(f"generating {len(items)} trajectories", flush=True)
    trajs, secs = generate(items, sizes, riemann_n=args.riemann_n, batch=args.batch,
                           log=lambda s: print(s, flush=True), source=args.families)
    if args.check_steps:
        rows = json.loads(Path(args.check_steps).read_text())["rows"]
        want = {(r[0], r[1]): r[3] for r in rows}
        got = {k: v["steps"] for k, v in trajs.items() if k in want}
        same = sum(got[k] == want[k] for k in got)
        print(f"step counts equal to the fixture's: {same} of {len(got)}", flush=True)
        for k in sorted(got):
            if got[k] != want[k]:
                print(f"  {k[0]} {k[1]}: {got[k]} steps, fixture {want[k]}", flush=True)
    if args.no_pack:
        raise SystemExit(0)
    print(f"packing into {out}", flush=True)
    ok = pack(trajs, sizes, out, log=lambda s: print(s, flush=True))
    print("ALL CHECKS PASSED" if ok else "SOME CHECKS FAILED: see the lines above", flush=True)
    raise SystemExit(0 if ok else 1)


if __name__ == "__main__":
    main()

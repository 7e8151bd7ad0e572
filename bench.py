"""Benchmark: FP64 cell-updates/s of the 2D Euler FV step (HLLC + van Leer MUSCL, SSP-RK2, CFL 0.5).

Workload: the random-Riemann batch -- 64 instances of 512^2, periodic, each seeded by its index,
four constant quadrants split at a random corner -- advanced in the device-resident loop (per-
instance dt fused into the stage-2 kernel, RK2 steps captured in a hipGraph).  One "step" advances
every instance of the batch by one SSP-RK2 step.  Inputs are resident in HBM before timing starts.

The cost of a step depends on the solution: the first few hundred steps, over nearly uniform
quadrants, run about 10% faster than steps over the developed flow.  The default K = 3808 is the
mean number of steps these 64 instances take to the workload's end time t = 0.5 (1931-7103 each),
and every instance takes exactly K timed steps (no end time stops one early), so the timed steps
sample the states of a whole trajectory.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--cpu-seconds S] [--no-learned]

Prints one JSON line.  roofline: the stage-2 kernel (the dominant launch), whose algorithmic HBM
traffic is 96 B per cell (reads U and U1, writes U^{n+1}); its average duration is measured with
HIP events on the handle's stream, over a second pass through the same steps (stage_times).
cpu_baseline: the reference solver compiled from its sources (oracle/_ref) when that build is
present, else the repo's CPU port of it, timed on the host cores on instance 0 of the batch.
learned_matched_l1: the metric's second half, the learned 64^2 solver's time-to-matched-L1 speed-up
over the plain solver (learned_eval).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "neural-pde-surrogate_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

N_INST, N = 64, 512
T_OPEN = 100.0  # an end time no instance reaches within the benchmark's steps
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
BYTES_STAGE1, BYTES_STAGE2 = 64, 96  # algorithmic HBM bytes per cell per stage (binary64, 4 variables)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(U0, seconds, threads=None):
    """The reference solver (or its port) on instance 0 of the batch for about `seconds`, with `threads`
    OpenMP threads (default: all the host's)."""
    import numpy as np
    import oracle2d
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    g = oracle2d.Grid.make(N, N, bc_x="periodic", bc_y="periodic")
    if oracle2d.have_reference():
        kind, runner = "reference", oracle2d.Reference()
        run = lambda U, k: runner.run(g, "hllc", "vanleer", 0.5, U, 0.5, nout=1, max_steps=k, threads=threads)
    else:
        kind, runner = "port", oracle2d.Oracle()
        run = lambda U, k: runner.run(g, "hllc", "vanleer", 0.5, U, 0.5, nout=1, max_steps=k, threads=threads)
    steps, elapsed = 4, 0.0
    while True:
        U = np.ascontiguousarray(U0.copy())
        t0 = time.perf_counter()
        r = run(U, steps)
        elapsed = time.perf_counter() - t0
        done = r["steps"]
        if elapsed >= seconds or steps >= 4096:
            break
        steps = int(steps * max(2.0, min(8.0, 1.2 * seconds / max(elapsed, 1e-3))))
    value = N * N * done / elapsed
    return {"value": value, "unit": "cell-updates/s", "cores": threads, "kind": kind,
            "sample": f"instance 0 of the batch (512^2 random Riemann, HLLC + van Leer), {done} SSP-RK2 steps "
                      f"from t = 0 in {elapsed:.2f} s on {threads} OpenMP threads of {cpu_model()}"}


LIB = ROOT / "neural-pde-surrogate_amd" / "lib" / "libnpde.so"
DRIVER = ROOT / "neural-pde-surrogate_amd" / "lib" / "npde-euler2d"
REF_CLI = ROOT / "oracle" / "_ref" / "euler2d"
COUNTERS = ROOT / "profiles" / "stage_kernels_64x512_counters.md"


def library_sha16():
    import hashlib
    return hashlib.sha256(LIB.read_bytes()).hexdigest()[:16] if LIB.exists() else None


def counters_stale():
    """True when the committed counter summary was profiled on another build of libnpde.so than the one
    loaded here (its `library_sha16 = ...` line differs), so its per-launch figures may not describe the
    kernels that ran."""
    if not COUNTERS.exists():
        return True
    for line in COUNTERS.read_text().splitlines():
        if line.startswith("library_sha16 = "):
            return line.split("=", 1)[1].strip() != library_sha16()
    return True


def done_line(out):
    """(steps, million cell-updates/s) from an euler2d-style "Done:" line (main2d.cpp:259-260)."""
    import re
    m = re.search(r"Done: (\d+) steps in [0-9.]+ s \(([0-9.]+) million cell-updates per second\)", out)
    return (int(m.group(1)), float(m.group(2))) if m else (None, None)


def cli_run(exe, args, threads=None):
    import subprocess
    import tempfile
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as d:
        p = subprocess.run([str(exe), *args, "--out", d], capture_output=True, text=True, env=env, timeout=300)
    steps, mcells = done_line(p.stdout)
    return {"rc": p.returncode, "steps": steps, "value": None if mcells is None else mcells * 1e6,
            "unit": "cell-updates/s"}


def timed_trajectory(b, ics, t_end, nout, sizes, upload=False, download=False):
    """Seconds of a recorded run of the batch from the states ics after a short untimed run with the same
    frame count and sizes (buffers and captured graphs kept): the device-resident loop alone, or with the
    state uploads before it and the trajectory downloads after it (end to end)."""
    import torch
    for k, U in enumerate(ics):
        b.set_state(k, U)
    b.run_trajectory(t_end / 50, nout, sizes=sizes)
    if not upload:
        for k, U in enumerate(ics):
            b.set_state(k, U)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if upload:
        for k, U in enumerate(ics):
            b.set_state(k, U)
    b.run_trajectory(t_end, nout, sizes=sizes)
    if download:
        for k in range(len(ics)):
            for m in sizes:
                b.trajectory(k, m)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def config_runs(cpu=True):
    """The benchmark's other workloads, each at its stated size, as cell-update rates (one cell-update = one
    cell advanced one SSP-RK2 step, main2d.cpp:258):
      lax_liu3_256   the reference's own CPU case (Lax-Liu configuration 3, 256^2, t = 0.3, 10 snapshots;
                     682 steps, experiments/results.md:92-94): the device loop in this process, the drop-in
                     driver npde-euler2d end to end (snapshot files included, its "Done:" line), and the
                     reference CLI built from its sources on all host threads and on one
      riemann_64x512 the benchmark batch end to end: 64 random-Riemann instances of 512^2 to t = 0.5 with 50
                     snapshots at 128^2 and 64^2 and each instance's own tail (uploads and downloads
                     included), and its device loop alone
      blast_96       the uneven batch: 96 blast instances of 128^2-1024^2 (4:3:2:1) to t = 0.2 with 50
                     snapshots at 128^2 and 64^2
      kh_1024        one held-out Kelvin-Helmholtz instance at 1024^2 (600 steps from t = 0)"""
    import numpy as np
    import npde
    from npde.families import blast, blast_size, kh, riemann
    out = {}
    g, t_end, U0 = npde.builtin_problem("lax_liu3", 256)
    with npde.Batch([g], "hllc", "vanleer", 0.5) as b:
        secs = timed_trajectory(b, [U0], t_end, 10, ())
        steps = int(b.time(0)[1])
    ll = {"steps": steps, "device_loop": {"value": 256 * 256 * steps / secs, "seconds": secs}}
    if DRIVER.exists():
        ll["driver_end_to_end"] = cli_run(DRIVER, ["--test", "lax_liu3", "--n", "256"])
    if cpu and REF_CLI.exists():
        cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        ll["reference_cli"] = dict(cli_run(REF_CLI, ["--test", "lax_liu3", "--n", "256"], threads=cores), cores=cores)
        ll["reference_cli_1_thread"] = cli_run(REF_CLI, ["--test", "lax_liu3", "--n", "256"], threads=1)
    out["lax_liu3_256"] = ll
    log(f"lax_liu3 256^2: {steps} steps, device loop {ll['device_loop']['value'] / 1e9:.2f} G cell-updates/s")

    n, B = 512, 64
    ics = [riemann(k, n)[0] for k in range(B)]
    with npde.Batch([npde.make_grid(n, n, bc_x=1, bc_y=1)] * B, "hllc", "vanleer", 0.5) as b:
        e2e = timed_trajectory(b, ics, 0.5, 50, (128, 64), upload=True, download=True)
        cu = float(np.sum(b.times()[1])) * n * n
        dev = timed_trajectory(b, ics, 0.5, 50, (128, 64))
    out["riemann_64x512"] = {"cell_updates": cu, "end_to_end": {"value": cu / e2e, "seconds": e2e},
                             "device_loop": {"value": cu / dev, "seconds": dev}}
    log(f"riemann 64 x 512^2 to t = 0.5: end to end {cu / e2e / 1e9:.2f} G/s, device loop {cu / dev / 1e9:.2f} G/s")

    Bb = 96
    sizes = [blast_size(k) for k in range(Bb)]
    ics = [blast(k)[0] for k in range(Bb)]
    with npde.Batch([npde.make_grid(m, m, bc_x=1, bc_y=1) for m in sizes], "hllc", "vanleer", 0.5) as b:
        secs = timed_trajectory(b, ics, 0.2, 50, (128, 64))
        st = b.times()[1]
        cu = float(sum(m * m * int(k) for m, k in zip(sizes, st)))
    out["blast_96"] = {"cell_updates": cu, "value": cu / secs, "seconds": secs,
                       "steps_min": int(st.min()), "steps_max": int(st.max()),
                       "sizes": {str(m): sizes.count(m) for m in (128, 256, 512, 1024)}}
    log(f"blast 96 instances to t = 0.2: {cu / secs / 1e9:.2f} G/s in {secs:.2f} s")

    import torch
    U, _, _ = kh(0, 1024)
    with npde.Batch([npde.make_grid(1024, 1024, bc_x=1, bc_y=1)], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, U)
        b.advance_steps(32, 2.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.advance_steps(600, 2.0)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
    out["kh_1024"] = {"value": 1024 * 1024 * 600 / secs, "seconds": secs, "steps": 600}
    log(f"kh 1024^2 single instance: {out['kh_1024']['value'] / 1e9:.2f} G/s")
    return out


def counter(name, layout):
    """A per-launch stage-2 figure for this task layout from the committed rocprofv3 counter summary
    (lines `stage2_<name> (<layout>) = value`), or None without a measurement of that layout.  The
    HBM bytes are FETCH_SIZE x 2 + WRITE_SIZE x 1, the factors a plain device copy of known size gives."""
    f = COUNTERS
    if not f.exists():
        return None
    key = f"stage2_{name} ({layout}) = "
    for line in f.read_text().splitlines():
        if line.startswith(key):
            return float(line[len(key):])
    return None


def valu_roofline(layout, launch_ms):
    """Stage 2 against the FP64 VALU: the profiled instruction counts per launch over the live launch
    time.  Issue peak: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD per 4 clocks at 2.4 GHz;
    the counts weigh the quarter-rate FP64 transcendentals 4x (SQ_ACTIVE_INST_VALU)."""
    active = counter("valu_active_per_launch", layout)
    if active is None:
        return None
    peak = 256 * 4 * 2.4e9 / 4
    rate = active / (launch_ms * 1e-3)
    out = {"bound": "fp64 valu issue", "achieved": rate, "peak": peak, "unit": "wave-instr/s", "frac": rate / peak,
           "insts_per_launch": counter("valu_insts_per_launch", layout), "active_per_launch": active,
           "counters_stale": counters_stale()}
    flops = counter("fp64_flops_per_launch", layout)
    if flops is not None:
        out["fp64_tflops"] = flops / (launch_ms * 1e-3) / 1e12
        out["fp64_peak_tflops"] = FP64_PEAK_TFLOPS
    return out


def learned_eval():
    """The metric's second half, time-to-matched-L1 on the random-Riemann family
    (npde.learned.pipeline.evaluate, experiments/solver_baseline.py:96-124): the committed flux
    corrector on the 64^2 solver against the plain solver at 64^2, 128^2 and 256^2, on 64 held-out
    instances (indices 100000-100063, never trained on), density L1 over 50 frames against the 512^2
    solve block-averaged to 64^2; the matched plain resolution and its wall time by log-log
    interpolation.  The learned wall time includes the fallback fine solves."""
    from npde.learned import pipeline as P
    from npde.learned.model import load_trained
    idx = list(range(100000, 100064))
    quiet = lambda *a: None
    truth, _ = P.generate(idx, log=quiet)
    res = P.evaluate(load_trained(), idx, truth[P.COARSE], log=quiet)
    lr, m = res["learned64"], res["matched"]
    log(f"learned 64^2: L1(rho) {lr['l1_rho']:.4f} in {lr['seconds']:.3f} s ({lr['fallbacks']} fallbacks, "
        f"{lr['fallback_seconds']:.3f} s of it); matched plain "
        f"{m['plain_resolution']:.0f}^2 in {m['plain_seconds']:.3f} s: {m['speedup']:.2f}x")
    return {"metric": "time-to-matched-L1 speed-up, learned 64^2 vs plain", "value": m["speedup"],
            "target": 5.0, "higher_is_better": True, "learned_l1_rho": lr["l1_rho"], "learned_seconds": lr["seconds"],
            "fallbacks": lr["fallbacks"], "fallback_seconds": lr["fallback_seconds"],
            "network_switched_off": lr["network_switched_off"],
            "matched_plain_resolution": m["plain_resolution"],
            "matched_plain_seconds": m["plain_seconds"],
            "plain": {str(n): v for n, v in res["plain"].items()},
            "sample": "64 held-out random-Riemann instances, t = 0 to 0.5, 50 frames, truth 512^2"}


def stage_times(b, ics, warmup, steps):
    """Average stage-1 and stage-2 launch durations over the same states as the timed steps: the batch
    restarts from its initial states, takes the warm-up steps, then the timed steps again one at a time
    with HIP events around each stage kernel (a step's cost depends on the flow it advances)."""
    for k, U in enumerate(ics):
        b.set_state(k, U)
    b.advance_steps(warmup, T_OPEN)
    return b.time_stages(steps)


def fast_variant(b, ics, steps, warmup):
    """The opt-in fast-arithmetic stage kernels on the same batch and launch shape: throughput, stage
    times and the largest deviation from the exact kernels after 32 steps from the initial states
    (relative to each variable's maximum magnitude).  Reported beside the exact figures, never in
    place of them."""
    import numpy as np
    import torch
    N_CHECK, CHECK_STEPS = 4, 32
    b.set_arithmetic("fast")
    for k, U in enumerate(ics):
        b.set_state(k, U)
    b.advance_steps(warmup, T_OPEN)
    s0 = [b.time(k)[1] for k in range(N_INST)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.advance_steps(steps, T_OPEN)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s1 = [b.time(k)[1] for k in range(N_INST)]
    value = sum(N * N * (c1 - c0) for c0, c1 in zip(s0, s1)) / wall
    ms1, ms2 = stage_times(b, ics, warmup, steps)
    runs = {}
    for mode in ("fast", "exact"):
        b.set_arithmetic(mode)
        for k, U in enumerate(ics):
            b.set_state(k, U)
        b.advance_steps(CHECK_STEPS, 0.5)
        runs[mode] = [b.get_state(k) for k in range(N_CHECK)]
    dev = 0.0
    for f, e in zip(runs["fast"], runs["exact"]):
        scale = np.abs(e).reshape(4, -1).max(axis=1).reshape(4, 1, 1)
        dev = max(dev, float(np.max(np.abs(f - e) / scale)))
    log(f"fast arithmetic: {value / 1e9:.2f} G cell-updates/s, stage1 {ms1:.3f} ms, stage2 {ms2:.3f} ms, "
        f"max deviation from exact {dev:.2e} after {CHECK_STEPS} steps")
    return {"value": value, "unit": "cell-updates/s", "ms_per_step": wall * 1e3 / steps, "stage1_launch_ms": ms1,
            "stage2_launch_ms": ms2, "bit_exact": False, "max_rel_deviation_from_exact": dev,
            "deviation_sample": f"instances 0-{N_CHECK - 1} after {CHECK_STEPS} steps from the initial states, "
                                f"max |fast - exact| / max|exact| per variable"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # multiples of the graph length (32 steps), so no captured launch runs idle
    ap.add_argument("--steps", type=int, default=3808)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-fast", action="store_true", help="skip the fast-arithmetic (not bit-exact) variant")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-solver time-to-matched-L1 evaluation")
    ap.add_argument("--no-configs", action="store_true", help="skip the other workloads (config_runs)")
    ap.add_argument("--rows", type=int, default=None, help="fix the band height (skips autotuning)")
    ap.add_argument("--wpb", type=int, default=None, help="fix the waves per block (skips autotuning)")
    ap.add_argument("--row-groups", type=int, default=None, choices=(0, 1, 2),
                    help="fix the row-group mode (skips autotuning)")
    ap.add_argument("--seams", type=int, default=None, choices=(0, 1, 2),
                    help="fix the seam mode: 64-column strips that solve their seams first (skips autotuning)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run a plain 1 GiB device copy (calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the batch would shard across GPUs "
              "by instances with no communication; the multi-GPU launch is not implemented.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import npde
    from npde.families import riemann

    torch.cuda.set_device(0)
    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 3 x 1 GiB device copies done")
        del src, dst
    grids = [npde.make_grid(N, N, bc_x=npde.PERIODIC, bc_y=npde.PERIODIC)] * N_INST
    ics = [riemann(k, N)[0] for k in range(N_INST)]
    b = npde.Batch(grids, "hllc", "vanleer", 0.5)
    for k, U in enumerate(ics):
        b.set_state(k, U)
    fixed = {k: v for k, v in (("rows_per_task", args.rows), ("waves_per_block", args.wpb),
                                ("row_groups", args.row_groups), ("seams", args.seams)) if v is not None}
    if fixed:
        b.set_tuning(**fixed)
        tuning = b.tuning
    else:
        tuning = b.tuning if args.no_tune else b.autotune(None)
    for k, U in enumerate(ics):  # autotuning restores the state; set it again anyway
        b.set_state(k, U)
    layout = b.task_layout
    log(f"tuning: {tuning}, task layout {layout}")

    b.advance_steps(args.warmup, T_OPEN)
    s0 = [b.time(k)[1] for k in range(N_INST)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.advance_steps(args.steps, T_OPEN)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s1 = [b.time(k)[1] for k in range(N_INST)]
    updates = sum(N * N * (c1 - c0) for c0, c1 in zip(s0, s1))
    assert updates == N_INST * N * N * args.steps, "an instance stopped before taking every timed step"
    value = updates / wall
    ms_step = wall * 1e3 / args.steps

    ms1, ms2 = stage_times(b, ics, args.warmup, args.steps)
    cells = N_INST * N * N
    ach2 = BYTES_STAGE2 * cells / (ms2 * 1e-3) / 1e9
    ach1 = BYTES_STAGE1 * cells / (ms1 * 1e-3) / 1e9
    log(f"stage1 {ms1:.3f} ms ({ach1:.0f} GB/s), stage2 {ms2:.3f} ms ({ach2:.0f} GB/s), "
        f"step {ms_step:.3f} ms, {value / 1e9:.2f} G cell-updates/s")
    fast = None if args.no_fast else fast_variant(b, ics, args.steps, args.warmup)
    b.close()
    learned = None if args.no_learned else learned_eval()
    configs = None if args.no_configs else config_runs(cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ics[0], args.cpu_seconds)
        one = cpu_baseline(ics[0], args.cpu_seconds / 2, threads=1)
        cpu["single_thread"] = {"value": one["value"], "sample": one["sample"]}
        log(f"cpu baseline ({cpu['kind']}): {cpu['value'] / 1e6:.1f} M cell-updates/s on {cpu['cores']} threads, "
            f"{one['value'] / 1e6:.2f} M on 1")
    out = {
        "metric": "FP64 cell-updates/s (HLLC+vanLeer)", "value": value, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "random-Riemann batch: 64 instances x 512^2, periodic, HLLC + van Leer MUSCL, "
                               "SSP-RK2, CFL 0.5, per-instance dt, device-resident loop (hipGraph); every "
                               "instance takes the timed steps from t = 0 (default: the mean step count to t = 0.5)",
                   "instances": N_INST, "n": N, "tuning": tuning},
        "roofline": {"bound": "hbm", "kernel": f"npde::exact::k_stage<HLLC, vanLeer, 2, false, {layout}>",
                     "achieved": ach2, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach2 / HBM_PEAK_GBS,
                     "traffic": counter("hbm_bytes_per_launch", layout), "library_sha16": library_sha16(),
                     "traffic_stale": counters_stale(), "bytes_per_cell": BYTES_STAGE2, "launch_ms": ms2,
                     "stage1_launch_ms": ms1, "stage1_achieved": ach1,
                     # the whole step: every cell-update's algorithmic 160 B (stage 1 reads U and writes
                     # U1, stage 2 reads U and U1 and writes U) at the measured cell-update rate
                     "step_bytes_per_cell_update": BYTES_STAGE1 + BYTES_STAGE2,
                     "step_achieved": value * (BYTES_STAGE1 + BYTES_STAGE2) / 1e9,
                     "step_frac": value * (BYTES_STAGE1 + BYTES_STAGE2) / 1e9 / HBM_PEAK_GBS},
        "valu_roofline": valu_roofline(layout, ms2),
        "cpu_baseline": cpu,
        "fast_arith_variant": fast,
        "learned_matched_l1": learned,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

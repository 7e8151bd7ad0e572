"""GPU parity at the benchmark workloads' own sizes and horizons (the HIP path through its C-ABI against the
oracle): whole runs rather than the first few steps, the held-out Kelvin-Helmholtz family at 1024^2 and
128^2, the fast-arithmetic variant's 100-step bound, and the state, tuning and cache calls around a run.

Bar: bit-for-bit equality of every conserved value, frame and step count (the reference solver is binary64
without contraction, and so are the kernels); the fast variant's bound is stated where it is tested.
"""
import time

import numpy as np
import pytest

from conftest import bitwise_equal

import oracle2d

pytestmark = pytest.mark.gpu

same = bitwise_equal


def periodic(n):
    return oracle2d.Grid.make(n, n, bc_x="periodic", bc_y="periodic")


def totals_drift(U, U0):
    """max over the four conserved variables of |sum(U) - sum(U0)| / sum|U0| (sum|U0| > 0 also where a
    momentum total is near zero)."""
    d = np.abs(U.reshape(4, -1).sum(axis=1) - U0.reshape(4, -1).sum(axis=1))
    return float((d / np.abs(U0).reshape(4, -1).sum(axis=1)).max())


def test_lax_liu_256_full_run_matches_oracle(npde_mod, oracle):
    """The reference's own CPU case (Lax-Liu configuration 3, 256^2, HLLC + van Leer, transmissive, to
    t = 0.3 with the driver's default 10 snapshots): 682 steps (experiments/results.md:94), the same
    snapshot times and the final state bit for bit."""
    g, t_end, U0 = npde_mod.builtin_problem("lax_liu3", 256)
    go, _, Uo = oracle.setup("lax_liu3", 256)
    assert same(U0, Uo)
    with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, U0)
        b.run_trajectory(t_end, 10, sizes=(64,))
        t, steps = b.time(0)
        U, ft, fr = b.get_state(0), b.frame_times(0), b.trajectory(0, 64)
    ro = oracle.run(go, "hllc", "vanleer", 0.5, Uo, t_end, nout=10, sizes=(64,))
    assert steps == ro["steps"] == 682
    assert t == ro["t"] == t_end
    assert same(ft, ro["frame_t"])
    assert same(fr, ro["frames"][64])
    assert same(U, Uo)


def test_riemann_batch_instance_matches_oracle_over_the_whole_trajectory(npde_mod, oracle):
    """The benchmark batch end to end (64 random-Riemann instances of 512^2, periodic, each with its own dt,
    to t = 0.5 with 50 snapshots block-averaged to 128^2 and 64^2): instance 0 equals the oracle's run over
    the whole horizon -- step count, snapshot times, both snapshot levels and the final state bit for
    bit -- and every instance conserves its periodic totals to rounding."""
    from npde.families import riemann
    n, B = 512, 64
    ics = [riemann(k, n)[0] for k in range(B)]
    with npde_mod.Batch([npde_mod.make_grid(n, n, bc_x=1, bc_y=1)] * B, "hllc", "vanleer", 0.5) as b:
        for k, U in enumerate(ics):
            b.set_state(k, U)
        b.run_trajectory(0.5, 50, sizes=(128, 64))
        t, steps = b.time(0)
        U = b.get_state(0)
        ft = b.frame_times(0)
        fr = {m: b.trajectory(0, m) for m in (128, 64)}
        drift = max(totals_drift(b.get_state(k), ics[k]) for k in range(0, B, 7))
        all_steps = b.times()[1]
    a = ics[0].copy()
    ro = oracle.run(periodic(n), "hllc", "vanleer", 0.5, a, 0.5, nout=50, sizes=(128, 64))
    assert steps == ro["steps"] and t == ro["t"] == 0.5
    assert same(ft, ro["frame_t"])
    for m in (128, 64):
        assert same(fr[m], ro["frames"][m]), m
    assert same(U, a)
    assert drift < 1e-12, drift
    assert all_steps.min() > 1000  # every instance ran its own whole horizon


FAST_TOL_100_STEPS = 1e-12  # max |fast - exact| / max|exact| per variable after 100 steps


def test_fast_variant_within_bound_after_100_steps_on_the_benchmark_batch(npde_mod):
    """The fast-arithmetic variant (reciprocal divisions, shortened square root, contraction) on the
    benchmark batch (64 x 512^2 random Riemann, the exact run's launch shape): after 100 CFL steps from the
    initial states every instance is within 1e-12 of the exact kernels, relative to each variable's
    largest magnitude -- the non-bit-exact bound of the task -- and the two are not identical."""
    from npde.families import riemann
    n, B, STEPS = 512, 64, 100
    ics = [riemann(k, n)[0] for k in range(B)]
    out = {}
    with npde_mod.Batch([npde_mod.make_grid(n, n, bc_x=1, bc_y=1)] * B, "hllc", "vanleer", 0.5) as b:
        for mode in ("exact", "fast"):
            b.set_arithmetic(mode)
            for k, U in enumerate(ics):
                b.set_state(k, U)
            b.advance_steps(STEPS, 0.5)
            assert (b.times()[1] == STEPS).all()
            out[mode] = [b.get_state(k) for k in range(B)]
    worst = 0.0
    for e, f in zip(out["exact"], out["fast"]):
        scale = np.abs(e).reshape(4, -1).max(axis=1).reshape(4, 1, 1)
        worst = max(worst, float(np.max(np.abs(f - e) / scale)))
    assert 0.0 < worst <= FAST_TOL_100_STEPS, worst


def test_kh_family_at_1024_matches_oracle_and_conserves(npde_mod, oracle):
    """The held-out Kelvin-Helmholtz family at its plain-solver size (instance 0 at 1024^2, periodic): 50
    steps equal the oracle's bit for bit, and the periodic totals of all four conserved variables hold to
    rounding."""
    from npde.families import kh
    n, STEPS = 1024, 50
    U0, _, _ = kh(0, n)
    with npde_mod.Batch([npde_mod.make_grid(n, n, bc_x=1, bc_y=1)], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, U0)
        b.advance_steps(STEPS, 2.0)
        t, steps = b.time(0)
        U = b.get_state(0)
    a = U0.copy()
    ro = oracle.run(periodic(n), "hllc", "vanleer", 0.5, a, 2.0, max_steps=STEPS)
    assert steps == ro["steps"] == STEPS and t == ro["t"]
    assert same(U, a)
    assert totals_drift(U, U0) < 1e-12


def test_kh_coarse_run_from_the_fine_state_matches_oracle(npde_mod, oracle):
    """The held-out protocol's plain coarse solve: the 1024^2 KH instance advanced on the GPU to its start
    time t_s in [0.5, 1.5], block-averaged in binary64 to 128^2, then 50 snapshot intervals to t = 2 at
    128^2 -- equal to the oracle's run from the same coarse state (step count, snapshot times, snapshots
    and final state bit for bit)."""
    from npde.families import kh
    from npde.learned.pipeline import block_average_f64
    n_fine, n = 1024, 128
    U0, t_end, meta = kh(0, n_fine)
    ts = float(meta["t_start"])
    assert 0.5 <= ts <= 1.5
    with npde_mod.Batch([npde_mod.make_grid(n_fine, n_fine, bc_x=1, bc_y=1)], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, U0)
        b.advance_to(ts)
        assert b.time(0)[0] == ts
        Uc = block_average_f64(b.get_state(0), n)
    with npde_mod.Batch([npde_mod.make_grid(n, n, bc_x=1, bc_y=1)], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, Uc)
        b.run_trajectory(t_end - ts, 50, sizes=(n,))
        t, steps = b.time(0)
        U, ft, fr = b.get_state(0), b.frame_times(0), b.trajectory(0, n)
    a = Uc.copy()
    ro = oracle.run(periodic(n), "hllc", "vanleer", 0.5, a, t_end - ts, nout=50, sizes=(n,))
    assert steps == ro["steps"] and t == ro["t"]
    assert same(ft, ro["frame_t"])
    assert same(fr, ro["frames"][n])
    assert same(U, a)


def test_autotune_mid_run_then_the_run_continues_exactly(npde_mod, oracle, tmp_path):
    """npde_autotune in the middle of a run (it times candidate launch shapes on the handle's own state and
    restores it): the run continues to the oracle's bits, step count and time.  The cache file round trip:
    a second handle of the same configuration reads the chosen tuning back without timing, and a line
    without the layout fields keeps the handle's own layout choices."""
    from npde.families import riemann
    n, B, T1, T2 = 256, 4, 0.04, 0.08
    ics = [riemann(10 + k, n)[0] for k in range(B)]
    grid = npde_mod.make_grid(n, n, bc_x=1, bc_y=1)
    cache = tmp_path / "tune.txt"
    with npde_mod.Batch([grid] * B, "hllc", "vanleer", 0.5) as b:
        for k, U in enumerate(ics):
            b.set_state(k, U)
        b.advance_to(T1)
        before = b.times()
        t0 = time.perf_counter()
        chosen = b.autotune(str(cache))
        tune_s = time.perf_counter() - t0
        assert b.tuning == chosen
        after = b.times()
        assert same(before[0], after[0]) and (before[1] == after[1]).all()  # run state restored
        b.advance_to(T2)
        got = [(b.time(k), b.get_state(k)) for k in range(B)]
    for k in range(B):
        a = ics[k].copy()
        ro = oracle.run(periodic(n), "hllc", "vanleer", 0.5, a, T2, nout=2)
        assert got[k][0] == (ro["t"], ro["steps"]), k
        assert same(got[k][1], a), k
    lines = cache.read_text().splitlines()
    assert len(lines) == 1
    key, fields = lines[0].split()[0], [int(v) for v in lines[0].split()[1:]]
    assert fields == [chosen["rows_per_task"], chosen["waves_per_block"], chosen["row_groups"], chosen["seams"],
                      chosen["waves_per_simd"]]
    with npde_mod.Batch([grid] * B, "hllc", "vanleer", 0.5) as b:
        for k, U in enumerate(ics):
            b.set_state(k, U)
        t0 = time.perf_counter()
        assert b.autotune(str(cache)) == chosen
        assert time.perf_counter() - t0 < 0.25 * tune_s  # read back, not timed again
        # a hand-written line with a band height the tuner never tries and no layout fields
        b.set_tuning(row_groups=1, seams=2)
        cache.write_text(f"{key} 7 2\n")
        t = b.autotune(str(cache))
        assert (t["rows_per_task"], t["waves_per_block"]) == (7, 2)
        assert (t["row_groups"], t["seams"]) == (1, 2) and b.task_layout == "seams"
        b.advance_to(T1)
        a = ics[1].copy()
        ro = oracle.run(periodic(n), "hllc", "vanleer", 0.5, a, T1)
        assert b.time(1) == (ro["t"], ro["steps"]) and same(b.get_state(1), a)
        # a line with an occupancy cap: read back and enforced, same bits
        cache.write_text(f"{key} 16 1 0 2 1\n")
        t = b.autotune(str(cache))
        assert (t["rows_per_task"], t["waves_per_block"], t["seams"], t["waves_per_simd"]) == (16, 1, 2, 1)
        assert b.occupancy <= 1
        for k, U in enumerate(ics):
            b.set_state(k, U)
        b.advance_to(T1)
        assert b.time(1) == (ro["t"], ro["steps"]) and same(b.get_state(1), a)


def test_device_state_calls_match_the_host_calls(npde_mod):
    """npde_set_state_device / npde_get_state_device: a state handed over from a torch tensor steps to the
    same bits as the same state set from host memory, and reads back into a torch tensor unchanged.  The
    source tensor is released and its memory overwritten right after the call (the copy has completed)."""
    import torch
    from npde.families import riemann
    n = 192
    U0 = riemann(5, n)[0]
    grid = npde_mod.make_grid(n, n, bc_x=1, bc_y=0)
    with npde_mod.Batch([grid, grid], "hllc", "vanleer", 0.5) as b:
        src = torch.from_numpy(U0).cuda()
        torch.cuda.synchronize()
        b.set_state_device(1, src.data_ptr())
        src.fill_(-1.0)  # reuse of the source after the call must not reach the handle's copy
        del src
        b.set_state(0, U0)
        b.advance_steps(20, 1.0)
        assert b.time(0) == b.time(1) and b.time(0)[1] == 20
        assert same(b.get_state(0), b.get_state(1))
        dst = torch.empty(4, n, n, dtype=torch.float64, device="cuda")
        b.get_state_device(1, dst.data_ptr())
        assert same(dst.cpu().numpy(), b.get_state(0))


def test_kh_learned_coarse_run_with_fallback_at_the_stated_sizes(npde_mod):
    """The held-out protocol at its own sizes (fine 1024^2, learned 128^2 from t_s, 50 intervals to t = 2),
    natural and with the fallback forced: every frame is finite and positive in density, a learned run
    keeps the flux form (periodic mass and energy of its frames hold to float32 rounding), and a run that
    falls back takes the fine solve's frames bit for bit and is charged its time."""
    import torch
    from npde.learned.model import load_trained
    from npde.learned.pipeline import evaluate_kh
    model = load_trained(device="cuda")
    runs = [evaluate_kh(model, [0], n=128, n_fine=1024, log=lambda *a: None, keep_frames=True),
            evaluate_kh(model, [0], n=128, n_fine=1024, log=lambda *a: None, keep_frames=True, force_fallback=True)]
    torch.cuda.synchronize()
    for res, forced in zip(runs, (False, True)):
        (r,) = res["instances"]
        fr = r["frames"]
        assert fr["learned"].shape == (51, 4, 128, 128)
        for v in fr.values():
            assert np.isfinite(v).all() and v[:, 0].min() > 0.0
        assert r["fallback"] or not forced
        if r["fallback"]:
            assert same(fr["learned"], fr["truth"]) and r["l1_rho_learned"] == 0.0
            assert r["learned_seconds"] >= r["fine_seconds"]
        else:
            tot = fr["learned"].astype(np.float64).sum(axis=(-1, -2))  # (51, 4)
            for q in (0, 3):
                assert np.abs(tot[:, q] - tot[0, q]).max() <= 1e-5 * abs(tot[0, q]), q
            assert r["l1_rho_learned"] > 0.0
        assert r["l1_rho_plain"] > 0.0

"""GPU parity: the HIP path (libnpde.so through its C-ABI) against the oracle and the golden vectors.

Bar: bit-for-bit equality (bit patterns, so +0 and -0 differ) of every conserved value, every dt and the step count -- the
reference solver is binary64 without contraction, and so are the kernels.
"""
import numpy as np
import pytest

from conftest import bitwise_equal, run_case_grid

import oracle2d

pytestmark = pytest.mark.gpu

SCHEMES = ["rusanov", "hllc"]
LIMS = ["none", "unlimited", "minmod", "vanleer"]
RUN_NAMES = ["lax_liu3_64_hllc_vanleer", "lax_liu3_48_rusanov_minmod", "sod_x_64_hllc_unlimited",
             "sod_y_64_rusanov_vanleer", "sine2d_32_hllc_none", "kh_32_rusanov_vanleer", "kh_40_hllc_minmod",
             "riemann_48_hllc_vanleer", "riemann_40_hllc_minmod"]


def same(a, b):
    return bitwise_equal(a, b)


def to_npde_grid(npde, g):
    return npde.make_grid(g.nx, g.ny, g.x_min, g.x_max, g.y_min, g.y_max, g.bc_x, g.bc_y)


@pytest.mark.parametrize("name", RUN_NAMES)
def test_golden_runs_device_loop(npde_mod, golden, name):
    print(f"case {name}", flush=True)
    r = golden("runs.npz")
    g = run_case_grid(r[f"{name}/grid"])
    _, scheme, lim = (str(x) for x in r[f"{name}/names"])
    t_end, nout = r[f"{name}/cfg"]
    with npde_mod.Batch([to_npde_grid(npde_mod, g)], scheme, lim, 0.5) as b:
        b.set_state(0, r[f"{name}/U0"])
        b.run_trajectory(float(t_end), int(nout), sizes=())
        t, steps = b.time(0)
        U = b.get_state(0)
        ft = b.frame_times(0)
    assert steps == int(r[f"{name}/steps"])
    assert t == float(r[f"{name}/t"])
    assert same(ft, r[f"{name}/frame_t"])
    assert same(U, r[f"{name}/U"])


@pytest.mark.parametrize("name", ["lax_liu3_64_hllc_vanleer", "sod_y_64_rusanov_vanleer", "riemann_40_hllc_minmod"])
def test_golden_runs_host_driven_steps(npde_mod, golden, name):
    """Solver.compute_dt / Solver.step, the reference's call pattern, reproduce the dt sequence."""
    r = golden("runs.npz")
    g = run_case_grid(r[f"{name}/grid"])
    _, scheme, lim = (str(x) for x in r[f"{name}/names"])
    dts = r[f"{name}/dts"]
    s = npde_mod.Solver(to_npde_grid(npde_mod, g), scheme, lim, 0.5)
    U = r[f"{name}/U0"].copy()
    t_end, nout = r[f"{name}/cfg"]
    t, k, step = 0.0, 1, 0
    while step < 25:
        t_next = 0.0 + (float(t_end) - 0.0) * k / int(nout)
        if not t < t_next:
            k += 1
            continue
        dt = min(s.compute_dt(U), t_next - t)
        assert dt == dts[step], step
        s.step(U, dt)
        t += dt
        step += 1
    s.close()


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("lim", LIMS)
def test_matrix_against_oracle(npde_mod, oracle, scheme, lim):
    """Every scheme x limiter on a non-square, mixed-BC grid whose width is not a multiple of the strip."""
    from npde.families import riemann
    print(f"case matrix {scheme} {lim}", flush=True)
    U0, _, _ = riemann(11, 70)
    U0 = np.ascontiguousarray(U0[:, :53, :])
    g = oracle2d.Grid.make(70, 53, y=(0.0, 53 / 70), bc_x="transmissive", bc_y="periodic")
    a = U0.copy()
    ro = oracle.run(g, scheme, lim, 0.5, a, 0.04, nout=1, record_dts=True)
    with npde_mod.Batch([to_npde_grid(npde_mod, g)], scheme, lim, 0.5) as b:
        b.set_state(0, U0)
        b.advance_to(0.04)
        t, steps = b.time(0)
        U = b.get_state(0)
    assert steps == ro["steps"] and t == ro["t"]
    assert same(U, a)


def test_trajectory_matches_reference_cli(npde_mod, golden):
    t = golden("traj_ll64.npz")
    g, t_end, U0 = npde_mod.builtin_problem("lax_liu3", 64)
    with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, U0)
        b.run_trajectory(t_end, 4, sizes=(64, 32, 16))
        assert same(b.frame_times(0), t["t"])
        for m in (64, 32, 16):
            assert same(b.trajectory(0, m), t[f"U_{m}"]), m


def test_builtin_problems_match_oracle(npde_mod, oracle):
    for test, n in [("sod_x", 50), ("sod_y", 30), ("lax_liu3", 40), ("sine2d", 32), ("kh", 48)]:
        g, t_end, U = npde_mod.builtin_problem(test, n)
        go, to, Uo = oracle.setup(test, n)
        assert (g.nx, g.ny, g.bc_x, g.bc_y) == (go.nx, go.ny, go.bc_x, go.bc_y) and t_end == to
        assert same(U, Uo), test


def test_uneven_batch_each_instance_matches_oracle(npde_mod, oracle):
    """Instances of different sizes and boundaries in one launch; each must equal its own oracle run."""
    from npde.families import blast, riemann
    cases = []
    for k, (nx, ny, bx, by) in enumerate([(64, 64, 1, 1), (128, 32, 1, 0), (37, 53, 0, 0), (96, 96, 1, 1),
                                          (61, 61, 0, 1), (120, 60, 1, 1)]):
        U, _, _ = (riemann(k, max(nx, ny)) if k % 2 == 0 else blast(k, max(nx, ny)))
        cases.append((oracle2d.Grid.make(nx, ny, y=(0.0, ny / nx), bc_x=bx, bc_y=by),
                      np.ascontiguousarray(U[:, :ny, :nx])))
    T = 0.02
    with npde_mod.Batch([to_npde_grid(npde_mod, g) for g, _ in cases], "hllc", "vanleer", 0.5) as b:
        for k, (_, U) in enumerate(cases):
            b.set_state(k, U)
        b.advance_to(T)
        for k, (g, U) in enumerate(cases):
            a = U.copy()
            ro = oracle.run(g, "hllc", "vanleer", 0.5, a, T)
            t, steps = b.time(k)
            assert steps == ro["steps"] and t == ro["t"], k
            assert same(b.get_state(k), a), k


@pytest.mark.parametrize("tuning", [{}, dict(rows_per_task=1, waves_per_block=1)])
def test_grids_narrower_than_the_stencil_match_oracle(npde_mod, oracle, tuning):
    """Grids of 1-5 cells per side (every ghost of the stencil wraps or clamps onto one or two interior
    cells), mixed periodic and transmissive, in one batch and one by one: each equals its oracle run
    (pinned to the reference build on the same cases by the oracle tests)."""
    from test_oracle import TINY_GRIDS, tiny_state
    cases = [(oracle2d.Grid.make(nx, ny, y=(0.0, ny / nx), bc_x=bx, bc_y=by), tiny_state(nx, ny, 7 * nx + ny))
             for nx, ny, bx, by in TINY_GRIDS]
    expect = []
    for g, U in cases:
        a = U.copy()
        ro = oracle.run(g, "hllc", "vanleer", 0.5, a, 0.05, nout=2)
        expect.append((ro, a))
    for batch in ([list(range(len(cases)))] + [[k] for k in range(len(cases))]):
        with npde_mod.Batch([to_npde_grid(npde_mod, cases[k][0]) for k in batch], "hllc", "vanleer", 0.5) as b:
            if tuning:
                b.set_tuning(**tuning)
            for j, k in enumerate(batch):
                b.set_state(j, cases[k][1])
            b.run_trajectory(0.05, 2, sizes=())
            for j, k in enumerate(batch):
                ro, a = expect[k]
                t, steps = b.time(j)
                assert steps == ro["steps"] and t == ro["t"], (batch, k)
                assert same(b.get_state(j), a), (batch, k)


def test_anisotropic_cells_match_oracle(npde_mod, oracle):
    """Cells with dy != dx on offset domains (the cases the oracle tests pin to the reference build),
    batched: each instance equals its oracle run."""
    from npde.families import riemann
    from test_oracle import ANISOTROPIC
    cases = [(oracle2d.Grid.make(nx, ny, x=xr, y=yr, bc_x=bx, bc_y=by),
              np.ascontiguousarray(riemann(11, max(nx, ny))[0][:, :ny, :nx])) for nx, ny, xr, yr, bx, by in ANISOTROPIC]
    with npde_mod.Batch([to_npde_grid(npde_mod, g) for g, _ in cases], "hllc", "vanleer", 0.5) as b:
        for k, (_, U) in enumerate(cases):
            b.set_state(k, U)
        b.run_trajectory(0.04, 2, sizes=())
        for k, (g, U) in enumerate(cases):
            a = U.copy()
            ro = oracle.run(g, "hllc", "vanleer", 0.5, a, 0.04, nout=2)
            t, steps = b.time(k)
            assert steps == ro["steps"] and t == ro["t"], k
            assert same(b.get_state(k), a), k


def random_case(rng):
    """One seeded configuration: grid (some widths seam- or row-group-eligible), boundaries, scheme,
    limiter, launch shape, and a four-quadrant initial state with random corner and states."""
    from npde.solver import to_conserved
    nx = int(rng.choice([1, 3, 17, 60, 64, 100, 128, 131]))
    ny = int(rng.choice([1, 2, 9, 33, 64, 96]))
    bx, by = int(rng.integers(0, 2)), int(rng.integers(0, 2))
    scheme = str(rng.choice(["rusanov", "hllc"]))
    lim = str(rng.choice(["none", "unlimited", "minmod", "vanleer"]))
    tuning = dict(rows_per_task=int(rng.choice([0, 1, 3, 8, 64])), waves_per_block=int(rng.choice([1, 2, 4])),
                  seams=int(rng.choice([0, 2])), row_groups=int(rng.choice([0, 1, 2])),
                  steps_per_graph=int(rng.choice([1, 4, 32])), use_graph=int(rng.choice([0, 1])))
    cx, cy = int(rng.integers(0, nx + 1)), int(rng.integers(0, ny + 1))
    W = np.stack([rng.uniform(0.1, 2.0, 4), rng.uniform(-1.0, 1.0, 4), rng.uniform(-1.0, 1.0, 4),
                  rng.uniform(0.1, 2.0, 4)], axis=1)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    quad = (ii >= cx).astype(int) + 2 * (jj >= cy).astype(int)
    U = np.ascontiguousarray(to_conserved(*(W[quad, q] for q in range(4))))
    g = oracle2d.Grid.make(nx, ny, y=(0.0, ny / nx), bc_x=bx, bc_y=by)
    return g, U, scheme, lim, tuning


@pytest.mark.parametrize("seed", range(64))
def test_random_configurations_match_oracle(npde_mod, oracle, seed):
    """Seeded random configurations over grid size, boundaries, scheme, limiter and launch shape (task
    layout, band height, workgroup size, graph length, graphs or stream launches): the GPU run equals the
    oracle's -- step count, time and bits, or the same first non-physical cell when the run fails."""
    g, U, scheme, lim, tuning = random_case(np.random.default_rng(1000 + seed))
    T = 0.03
    a = U.copy()
    ro = oracle.run(g, scheme, lim, 0.5, a, T, nout=3)
    with npde_mod.Batch([to_npde_grid(npde_mod, g)], scheme, lim, 0.5) as b:
        b.set_tuning(**tuning)
        b.set_state(0, U)
        if ro["bad"] is None:
            b.run_trajectory(T, 3, sizes=())
            t, steps = b.time(0)
            assert steps == ro["steps"] and t == ro["t"], (seed, b.task_layout)
            assert same(b.get_state(0), a), (seed, b.task_layout)
        else:
            with pytest.raises(npde_mod.NonPhysicalStateError) as e:
                b.run_trajectory(T, 3, sizes=())
            assert (e.value.i, e.value.j) == ro["bad"], seed
            assert b.time(0)[1] == ro["steps"], seed


@pytest.mark.parametrize("seed", range(8))
def test_random_uneven_batches_match_oracle(npde_mod, oracle, seed):
    """Seeded batches of 12 random instances (the grids, boundaries and states of random_case) sharing one
    scheme, limiter and launch shape: every instance equals its own oracle run -- step count, time and
    bits -- or, when one fails, the batch reports the oracle's first non-physical cell of the first
    failing instance."""
    rng = np.random.default_rng(5000 + seed)
    cases = [random_case(np.random.default_rng(int(rng.integers(1 << 30)))) for _ in range(12)]
    _, _, scheme, lim, tuning = cases[0]
    T = 0.03
    expect = []
    for g, U, _, _, _ in cases:
        a = U.copy()
        expect.append((oracle.run(g, scheme, lim, 0.5, a, T, nout=3), a))
    with npde_mod.Batch([to_npde_grid(npde_mod, c[0]) for c in cases], scheme, lim, 0.5) as b:
        b.set_tuning(**tuning)
        for k, c in enumerate(cases):
            b.set_state(k, c[1])
        failing = [k for k, (ro, _) in enumerate(expect) if ro["bad"] is not None]
        if not failing:
            b.run_trajectory(T, 3, sizes=())
            for k, (ro, a) in enumerate(expect):
                t, steps = b.time(k)
                assert steps == ro["steps"] and t == ro["t"], (seed, k, b.task_layout)
                assert same(b.get_state(k), a), (seed, k, b.task_layout)
        else:
            with pytest.raises(npde_mod.NonPhysicalStateError) as e:
                b.run_trajectory(T, 3, sizes=())
            # every instance runs on its own, so each failing one stops at its own first bad cell; the batch
            # reports the lowest-numbered of them
            assert e.value.instance == failing[0], (seed, e.value.instance, failing)
            assert (e.value.i, e.value.j) == expect[failing[0]][0]["bad"], seed


FAST_TOL_RANDOM = 1e-9  # max |fast - exact| / max|exact| per variable after the configuration's steps


@pytest.mark.parametrize("seed", range(24))
def test_random_configurations_fast_arithmetic_stays_close(npde_mod, oracle, seed):
    """The fast-arithmetic kernels on the seeded random configurations that the exact run completes:
    after as many CFL steps as the exact run takes, every conserved variable is within FAST_TOL_RANDOM of
    the exact result, relative to its largest magnitude."""
    g, U, scheme, lim, tuning = random_case(np.random.default_rng(1000 + seed))
    ro = oracle.run(g, scheme, lim, 0.5, U.copy(), 0.03, nout=3)
    if ro["bad"] is not None:
        return  # the exact run fails; the error path is covered by the exact test
    out = {}
    for mode in ("exact", "fast"):
        with npde_mod.Batch([to_npde_grid(npde_mod, g)], scheme, lim, 0.5) as b:
            b.set_tuning(**tuning)
            b.set_arithmetic(mode)
            b.set_state(0, U)
            b.advance_steps(ro["steps"], 1e3)
            assert b.time(0)[1] == ro["steps"]
            out[mode] = b.get_state(0)
    scale = np.abs(out["exact"]).reshape(4, -1).max(axis=1).reshape(4, 1, 1)
    dev = float(np.max(np.abs(out["fast"] - out["exact"]) / scale))
    assert dev <= FAST_TOL_RANDOM, (seed, dev)


def test_stage_timing_takes_the_device_loops_steps(npde_mod):
    """npde_time_stages advances every instance by its next CFL steps: the same time, step count and
    bits as npde_advance_steps over as many steps."""
    g, _, U0 = npde_mod.builtin_problem("lax_liu3", 64)
    with npde_mod.Batch([g, g], "hllc", "vanleer", 0.5) as a, npde_mod.Batch([g, g], "hllc", "vanleer", 0.5) as b:
        for h in (a, b):
            h.set_state(0, U0)
            h.set_state(1, U0)
        ms1, ms2 = a.time_stages(25)
        b.advance_steps(25, 1e3)
        assert ms1 > 0 and ms2 > 0
        for k in range(2):
            assert a.time(k) == b.time(k) and a.time(k)[1] == 25
            assert same(a.get_state(k), b.get_state(k))


@pytest.mark.parametrize("spg", [1, 8, 32])
def test_frames_between_frame_free_launches_match_oracle(npde_mod, oracle, spg):
    """Recording runs launch the steps before the next frame time without the frame kernels when every
    stepping instance is far enough from it; every frame, frame time, step count and final state still
    equals the oracle's, for one instance and for two of different sizes (different landing steps)."""
    from npde.families import riemann
    NF, T = 6, 0.06
    specs = [(oracle2d.Grid.make(64, 64, bc_x=1, bc_y=1), np.ascontiguousarray(riemann(3, 64)[0])),
             (oracle2d.Grid.make(96, 96, bc_x=1, bc_y=1), np.ascontiguousarray(riemann(4, 96)[0]))]
    for cases in (specs[:1], specs):
        with npde_mod.Batch([to_npde_grid(npde_mod, g) for g, _ in cases], "hllc", "vanleer", 0.5) as b:
            b.set_tuning(steps_per_graph=spg)
            for k, (_, U) in enumerate(cases):
                b.set_state(k, U)
            b.run_trajectory(T, NF, sizes=(16,))
            for k, (g, U) in enumerate(cases):
                a = U.copy()
                ro = oracle.run(g, "hllc", "vanleer", 0.5, a, T, nout=NF, sizes=(16,))
                t, steps = b.time(k)
                assert steps == ro["steps"] and t == ro["t"], (len(cases), k)
                assert np.array_equal(b.frame_times(k), ro["frame_t"]), (len(cases), k)
                assert np.array_equal(b.trajectory(k, 16), ro["frames"][16]), (len(cases), k)
                assert same(b.get_state(k), a), (len(cases), k)


@pytest.mark.parametrize("spg", [1, 3, 8, 32])
def test_short_device_loops_match_oracle(npde_mod, oracle, spg):
    """Many short npde_advance_to calls of a few steps each (as a learned rollout makes between network
    calls) land on the oracle's frame times with its step count and bits, whatever device-loop sizes the
    launches take (sized from the run state: 1, 2, 4, ... iterations up to steps_per_graph)."""
    g, _, U0 = npde_mod.builtin_problem("lax_liu3", 64)
    go, _, _ = oracle.setup("lax_liu3", 64)
    NF, T = 12, 0.06
    a = U0.copy()
    ro = oracle.run(go, "hllc", "vanleer", 0.5, a, T, nout=NF)
    with npde_mod.Batch([g, g], "hllc", "vanleer", 0.5) as b:
        b.set_tuning(steps_per_graph=spg)
        b.set_state(0, U0)
        b.set_state(1, U0)
        for k in range(1, NF + 1):
            b.advance_to(0.0 + (T - 0.0) * k / NF)
        t, steps = b.time(0)
        assert steps == ro["steps"] and t == ro["t"] == T
        assert same(b.get_state(0), a) and same(b.get_state(1), a)


@pytest.mark.parametrize("tuning", [dict(rows_per_task=1, waves_per_block=1), dict(rows_per_task=7, waves_per_block=2),
                                    dict(rows_per_task=64, waves_per_block=4, use_graph=0),
                                    dict(rows_per_task=200, waves_per_block=4, steps_per_graph=3)])
def test_launch_shape_does_not_change_bits(npde_mod, oracle, tuning):
    g, t_end, U0 = npde_mod.builtin_problem("lax_liu3", 96)
    go, _, _ = oracle.setup("lax_liu3", 96)
    a = U0.copy()
    ro = oracle.run(go, "hllc", "vanleer", 0.5, a, 0.05, nout=2)
    with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
        b.set_tuning(**tuning)
        b.set_state(0, U0)
        b.run_trajectory(0.05, 2, sizes=(48,))
        assert b.time(0)[1] == ro["steps"]
        assert same(b.get_state(0), a)


@pytest.mark.parametrize("tuning", [dict(waves_per_block=1, seams=0), dict(waves_per_block=2, seams=0),
                                    dict(waves_per_block=1, seams=2, rows_per_task=16),
                                    dict(waves_per_block=4, seams=2, rows_per_task=32)])
def test_occupancy_cap_limits_resident_waves_and_keeps_bits(npde_mod, oracle, tuning):
    """waves_per_simd caps the stage kernels' resident wavefronts by an LDS reservation: the occupancy calculator
    sees at most the cap, caps an LDS reservation cannot enforce are refused, and every cap computes the same bits."""
    g, t_end, U0 = npde_mod.builtin_problem("lax_liu3", 128)
    go, _, _ = oracle.setup("lax_liu3", 128)
    a = U0.copy()
    ro = oracle.run(go, "hllc", "vanleer", 0.5, a, 0.02, nout=2)
    with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
        b.set_tuning(waves_per_simd=0, **tuning)
        natural = b.occupancy
        print(f"{tuning}: {natural} wavefronts per SIMD uncapped", flush=True)
        assert 1 <= natural <= 8
        accepted = []
        for cap in range(1, natural):
            try:
                b.set_tuning(waves_per_simd=cap)
            except npde_mod.NpdeError as e:
                assert "occupancy cap" in str(e)
                continue
            assert b.occupancy <= cap
            accepted.append(cap)
            b.set_state(0, U0)
            b.run_trajectory(0.02, 2, sizes=(64,))
            assert b.time(0)[1] == ro["steps"]
            assert same(b.get_state(0), a)
        print(f"  caps enforced: {accepted}", flush=True)
        assert natural == 1 or accepted or tuning["waves_per_block"] == 4  # (4-wave groups: caps 1, 2 need > 64 KiB)
        b.set_tuning(waves_per_simd=0)
        assert b.occupancy == natural


def test_nonphysical_state_raises_with_first_bad_cell(npde_mod, golden):
    r = golden("runs.npz")
    g, _, U0 = npde_mod.builtin_problem("sod_x", 100)
    with npde_mod.Batch([g], "rusanov", "unlimited", 0.5) as b:
        b.set_state(0, U0)
        with pytest.raises(npde_mod.NonPhysicalStateError) as e:
            b.advance_to(0.2)
        assert (e.value.i, e.value.j) == tuple(int(v) for v in r["fail_sod_unlimited/bad"])
        assert b.time(0)[1] == int(r["fail_sod_unlimited/steps"])
        assert same(b.get_state(0), r["fail_sod_unlimited/U"])


def test_sod_transpose_is_exact(npde_mod):
    gx, t_end, Ux = npde_mod.builtin_problem("sod_x", 200)
    gy, _, Uy = npde_mod.builtin_problem("sod_y", 200)
    with npde_mod.Batch([gx, gy], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, Ux)
        b.set_state(1, Uy)
        b.advance_to(t_end)
        assert b.time(0)[1] == b.time(1)[1]
        X, Y = b.get_state(0), b.get_state(1)
    assert np.abs(X - np.transpose(Y, (0, 2, 1))[[0, 2, 1, 3]]).max() == 0.0
    assert (X[2] == 0.0).all()


def test_sod_transpose_is_exact_across_layouts(npde_mod):
    """sod_x at 256 x 4 runs in seam mode (width a multiple of 64), sod_y at 4 x 256 in strips: the
    transposed results are still exactly equal (test_solver2d.cpp:86-95's bar, across layouts)."""
    gx, t_end, Ux = npde_mod.builtin_problem("sod_x", 256)
    gy, _, Uy = npde_mod.builtin_problem("sod_y", 256)
    with npde_mod.Batch([gx, gy], "hllc", "vanleer", 0.5) as b:
        b.set_tuning(seams=2)
        assert b.task_layout == "strips"   # one batch runs one layout: sod_y's width 4 rules seams out
    X, Y = {}, {}
    for g, U, out in ((gx, Ux, X), (gy, Uy, Y)):
        with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
            b.set_tuning(seams=2)
            out["layout"] = b.task_layout
            b.set_state(0, U)
            b.advance_to(t_end)
            out["steps"], out["U"] = b.time(0)[1], b.get_state(0)
    assert (X["layout"], Y["layout"]) == ("seams", "strips")
    assert X["steps"] == Y["steps"]
    assert np.abs(X["U"] - np.transpose(Y["U"], (0, 2, 1))[[0, 2, 1, 3]]).max() == 0.0


@pytest.mark.parametrize("tuning", [{}, dict(rows_per_task=3, waves_per_block=2, seams=0)])
def test_reference_solver_invariants(npde_mod, tuning):
    """test_solver2d.cpp's invariants on the GPU kernels, with its grids, states and bars:
    a uniform moving state on a periodic 16 x 12 grid stays uniform to 1e-13 after t = 0.5 (:36-53);
    Sod along x on 200 x 4 and along y on 4 x 200 (:56-67) take the same step count (:71), every row of
    the x-run is identical (:73-78), its y-momentum stays exactly 0 (:80-84), the y-run is the x-run
    transposed with momenta swapped exactly (:86-95), mass and energy are conserved to 1e-13 (:97-101),
    and the plateau at x = 0.78 is within 1e-3 of the exact rho*, p*, u* (:103-108)."""
    from npde.solver import to_conserved, to_primitive

    def run(g, U, t_end):
        with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
            if tuning:
                b.set_tuning(**tuning)
            b.set_state(0, U)
            b.advance_to(t_end)
            return b.time(0)[1], b.get_state(0)

    g = npde_mod.make_grid(16, 12, 0.0, 1.0, 0.0, 1.0, bc_x=1, bc_y=1)
    U0 = np.broadcast_to(to_conserved(1.3, 0.7, -0.4, 2.0)[:, None, None], (4, 12, 16)).copy()
    _, U = run(g, U0, 0.5)
    assert np.abs(U - U0).max() < 1e-13

    N, M = 200, 4
    gx = npde_mod.make_grid(N, M, 0.0, 1.0, 0.0, M / N)
    gy = npde_mod.make_grid(M, N, 0.0, M / N, 0.0, 1.0)
    left, right = to_conserved(1.0, 0.0, 0.0, 1.0), to_conserved(0.125, 0.0, 0.0, 0.1)
    xc = (np.arange(N) + 0.5) * (1.0 / N)
    col = np.where(xc[None, :] < 0.5, left[:, None], right[:, None])           # (4, N)
    Ux0 = np.ascontiguousarray(np.broadcast_to(col[:, None, :], (4, M, N)))
    Uy0 = np.ascontiguousarray(np.broadcast_to(col[:, :, None], (4, N, M)))
    nsx, Ux = run(gx, Ux0, 0.2)
    nsy, Uy = run(gy, Uy0, 0.2)
    assert nsx == nsy
    assert (Ux[0, 1:] == Ux[0, :1]).all()
    assert np.abs(Ux[2]).max() == 0.0
    assert np.abs(Ux - np.transpose(Uy, (0, 2, 1))[[0, 2, 1, 3]]).max() == 0.0
    cell = (1.0 / N) * (M / N / M)
    for q in (0, 3):
        assert abs(Ux[q].sum() * cell - Ux0[q].sum() * cell) < 1e-13
    rho, u, _, p = to_primitive(Ux[:, 0, int(0.78 * N)])
    assert abs(rho - 0.26557) < 1e-3 and abs(p - 0.30313) < 1e-3 and abs(u - 0.92745) < 1e-3


def test_trajectory_file_invariants(npde_mod):
    """tests/check_trajectory.py's checks on a periodic 128^2 random-Riemann trajectory at four levels:
    shapes and equally spaced frame times (:45-49); every coarse level the block average of the finest to
    1e-6 relative (:51-55); totals equal across levels (:57-61) and constant in time (:62-67); density and
    pressure positive (:69-74); and the finest level equal to the states of a frame-by-frame run
    (:76-85; here exactly, as float32 of the binary64 state)."""
    from npde.families import riemann
    from npde.solver import to_primitive
    n, nout, t_end, sizes = 128, 8, 0.2, (128, 64, 32, 16)
    g = npde_mod.make_grid(n, n, bc_x=1, bc_y=1)
    U0 = riemann(21, n)[0]
    with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, U0)
        b.run_trajectory(t_end, nout, sizes=sizes)
        levels = {m: b.trajectory(0, m).astype(np.float64) for m in sizes}
        t = b.frame_times(0)
    assert np.allclose(t, np.linspace(0.0, t_end, nout + 1), rtol=0, atol=1e-12)
    fine = levels[n]
    for m in sizes:
        assert levels[m].shape == (nout + 1, 4, m, m)
        f = n // m
        avg = fine.reshape(nout + 1, 4, m, f, m, f).mean(axis=(3, 5))
        assert np.abs(avg - levels[m]).max() / np.abs(fine).max() < 1e-6, m
    totals = {m: levels[m].mean(axis=(-2, -1)) for m in sizes}
    for m in sizes:
        assert np.abs(totals[m] - totals[n]).max() / np.abs(totals[n]).max() < 1e-6, m
        rho, _, _, p = to_primitive(np.moveaxis(levels[m], 1, 0))
        assert (rho > 0).all() and (p > 0).all(), m
    drift = np.abs(totals[n] - totals[n][:1]).max(axis=0) / np.abs(totals[n][0]).clip(1e-12)
    assert (drift[[0, 3]] < 1e-6).all()
    with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, U0)
        for k in range(nout + 1):
            b.advance_to(0.0 + (t_end - 0.0) * k / nout)
            assert np.array_equal(b.get_state(0).astype(np.float32), levels[n][k].astype(np.float32)), k


def test_lax_liu_256_published_step_count(npde_mod):
    g, t_end, U0 = npde_mod.builtin_problem("lax_liu3", 256)
    with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, U0)
        b.run_trajectory(t_end, 10, sizes=())
        assert b.time(0)[1] == 682


def test_ghosted_layout_step(npde_mod, oracle):
    """Solver.step on the reference's ghosted array: interior = oracle step, ghosts = apply_bcs(U^n)."""
    g, _, U0 = npde_mod.builtin_problem("lax_liu3", 40)
    go, _, _ = oracle.setup("lax_liu3", 40)
    A = np.zeros((44, 44, 4))
    A[2:42, 2:42, :] = np.moveaxis(U0, 0, -1)
    s = npde_mod.Solver(g, "hllc", "vanleer", 0.5)
    dt = s.compute_dt(A)
    dto, _, _, _ = oracle.compute_dt(go, 0.5, U0)
    assert dt == dto
    before = npde_mod.apply_bcs(g, A.copy())
    s.step(A, dt)
    a = U0.copy()
    oracle.step(go, "hllc", "vanleer", a, dt)
    assert same(np.moveaxis(A[2:42, 2:42, :], -1, 0), a)
    ring = np.ones((44, 44), bool)
    ring[2:42, 2:42] = False
    assert same(A[ring], before[ring])
    s.close()


def test_full_size_batch_properties(npde_mod, oracle):
    """64 x 512^2 random-Riemann batch (the benchmark workload): a few device-loop steps, then
    (a) instance 0 equals the oracle bit for bit, (b) every instance equals its own single-instance
    run, (c) periodic totals are conserved to rounding."""
    from npde.families import riemann
    n, B, T = 512, 64, 0.004
    grids = [npde_mod.make_grid(n, n, bc_x=1, bc_y=1)] * B
    ics = [riemann(k, n)[0] for k in range(B)]
    with npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as b:
        for k in range(B):
            b.set_state(k, ics[k])
        b.advance_to(T)
        states = [b.get_state(k) for k in (0, 17, 63)]
        steps = [b.time(k)[1] for k in (0, 17, 63)]
        mass = [b.get_state(k)[0].sum() for k in range(0, B, 9)]
    go = oracle2d.Grid.make(n, n, bc_x="periodic", bc_y="periodic")
    a = ics[0].copy()
    ro = oracle.run(go, "hllc", "vanleer", 0.5, a, T)
    assert steps[0] == ro["steps"]
    assert same(states[0], a)
    for idx, k in enumerate((17, 63)):
        with npde_mod.Batch([grids[0]], "hllc", "vanleer", 0.5) as s:
            s.set_state(0, ics[k])
            s.advance_to(T)
            assert s.time(0)[1] == steps[idx + 1]
            assert same(s.get_state(0), states[idx + 1])
    for j, k in enumerate(range(0, B, 9)):
        m0 = ics[k][0].sum()
        assert abs(mass[j] - m0) <= 1e-12 * abs(m0)


def test_uneven_blast_batch_full_size(npde_mod, oracle):
    """The uneven workload: 96 blast/bubble instances of 128^2-1024^2 (weights 4:3:2:1), periodic, in
    one batch for a few steps.  Each instance of every size equals its own single-instance run bit for
    bit, a 128^2 instance equals the oracle, and periodic totals are conserved to rounding."""
    from npde.families import blast, blast_size
    B, STEPS = 96, 6
    sizes = [blast_size(k) for k in range(B)]
    ics = [blast(k)[0] for k in range(B)]
    grids = [npde_mod.make_grid(n, n, bc_x=1, bc_y=1) for n in sizes]
    with npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as b:
        for k in range(B):
            b.set_state(k, ics[k])
        b.advance_steps(STEPS, 0.2)
        firsts = [sizes.index(n) for n in (128, 256, 512, 1024)]
        got = {k: b.get_state(k) for k in firsts}
        assert all(b.time(k)[1] == STEPS for k in range(B))
        mass = [(b.get_state(k)[0].sum(), ics[k][0].sum()) for k in range(0, B, 7)]
    for k in firsts:
        with npde_mod.Batch([grids[k]], "hllc", "vanleer", 0.5) as s:
            s.set_state(0, ics[k])
            s.advance_steps(STEPS, 0.2)
            assert same(s.get_state(0), got[k]), sizes[k]
    k = firsts[0]
    go = oracle2d.Grid.make(128, 128, bc_x="periodic", bc_y="periodic")
    a = ics[k].copy()
    ro = oracle.run(go, "hllc", "vanleer", 0.5, a, 0.2, max_steps=STEPS)
    assert ro["steps"] == STEPS and same(got[k], a)
    for m1, m0 in mass:
        assert abs(m1 - m0) <= 1e-12 * abs(m0)


FAST_TOL_10_STEPS = 1e-12  # max |fast - exact| / max|exact| per variable after 10 steps (measured ~5e-16)


@pytest.mark.parametrize("n,tuning,layout", [(96, {}, "strips"), (128, dict(seams=2), "seams")])
def test_fast_arithmetic_is_opt_in_and_close_to_exact(npde_mod, oracle, n, tuning, layout):
    """NPDE_ARITH_FAST (reciprocal divisions, shortened square root, contraction) stays within
    rounding-level distance of the exact kernels over 10 Lax-Liu steps, in the strip and the seam
    layouts, and switching back to NPDE_ARITH_EXACT reproduces the reference bits."""
    g, t_end, U0 = npde_mod.builtin_problem("lax_liu3", n)
    go, _, _ = oracle.setup("lax_liu3", n)
    with npde_mod.Batch([g], "hllc", "vanleer", 0.5) as b:
        if tuning:
            b.set_tuning(**tuning)
        assert b.task_layout == layout
        assert b.arithmetic == "exact"
        b.set_arithmetic("fast")
        assert b.arithmetic == "fast"
        b.set_state(0, U0)
        b.advance_steps(10, t_end)
        fast = b.get_state(0)
        b.set_arithmetic("exact")
        b.set_state(0, U0)
        b.advance_steps(10, t_end)
        exact = b.get_state(0)
    a = U0.copy()
    oracle.run(go, "hllc", "vanleer", 0.5, a, t_end, max_steps=10)
    assert same(exact, a)
    scale = np.abs(exact).reshape(4, -1).max(axis=1).reshape(4, 1, 1)
    dev = float(np.max(np.abs(fast - exact) / scale))
    assert 0.0 < dev <= FAST_TOL_10_STEPS, dev


@pytest.mark.parametrize("tuning,layout", [({}, "strips"), (dict(row_groups=2), "row_groups"),
                                           (dict(row_groups=0, seams=2), "seams")])
def test_fast_arithmetic_conserves_periodic_totals(npde_mod, tuning, layout):
    """The fast kernels keep the flux form in every task layout: periodic totals of all four conserved
    variables are conserved to rounding."""
    from npde.families import riemann
    U0 = riemann(3, 128)[0]
    with npde_mod.Batch([npde_mod.make_grid(128, 128, bc_x=1, bc_y=1)], "hllc", "vanleer", 0.5) as b:
        if tuning:
            b.set_tuning(**tuning)
        assert b.task_layout == layout
        b.set_arithmetic("fast")
        b.set_state(0, U0)
        b.advance_steps(40, 0.5)
        U = b.get_state(0)
    for q in range(4):
        assert abs(U[q].sum() - U0[q].sum()) <= 1e-12 * np.abs(U0[q]).sum(), q


@pytest.mark.parametrize("nx,nys", [(64, [64, 40]), (128, [96, 128, 33]), (512, [512])])
def test_row_groups_match_oracle_and_strips(npde_mod, oracle, nx, nys):
    """Row-group mode (periodic x, width 64*W: one workgroup per band, edge columns exchanged through
    LDS) gives the oracle's bits, the same as the strip mode, for W = 1, 2 and 8 and ragged band
    heights; the learned-increment kernels too."""
    from npde.families import riemann
    T = 0.01
    cases = []
    for k, ny in enumerate(nys):
        U = riemann(k + 5, max(nx, ny))[0][:, :ny, :nx]
        cases.append((oracle2d.Grid.make(nx, ny, y=(0.0, ny / nx), bc_x="periodic", bc_y=k % 2), np.ascontiguousarray(U)))
    results = {}
    for rg, rows, seams in ((2, 0, 2), (2, 7, 2), (1, 5, 2), (0, 16, 0), (0, 11, 2)):
        with npde_mod.Batch([to_npde_grid(npde_mod, g) for g, _ in cases], "hllc", "vanleer", 0.5) as b:
            b.set_tuning(row_groups=rg, rows_per_task=rows, seams=seams)
            expect = "row_groups" if rg == 2 or (rg == 1 and nx == 64) else ("seams" if seams else "strips")
            assert b.task_layout == expect
            for k, (_, U) in enumerate(cases):
                b.set_state(k, U)
            b.advance_to(T)
            results[(rg, rows, seams)] = [(b.time(k), b.get_state(k)) for k in range(len(cases))]
    for k, (g, U) in enumerate(cases):
        a = U.copy()
        ro = oracle.run(g, "hllc", "vanleer", 0.5, a, T)
        for key, res in results.items():
            (t, steps), got = res[k]
            assert steps == ro["steps"] and t == ro["t"], (key, k)
            assert same(got, a), (key, k)


@pytest.mark.parametrize("scheme,lim", [("hllc", "vanleer"), ("rusanov", "minmod")])
def test_seams_match_oracle_and_strips(npde_mod, oracle, scheme, lim):
    """Seam mode (widths that are multiples of 64: 64-column strips, each wavefront solving its strip
    boundaries first) gives the oracle's bits, the same as 60-column strips, on one batch of mixed widths
    (64-256), periodic and transmissive x and y, and ragged band heights."""
    from npde.families import riemann
    T = 0.01
    shapes = [(64, 64, "transmissive", "periodic"), (128, 96, "periodic", "transmissive"),
              (192, 70, "transmissive", "transmissive"), (256, 128, "periodic", "periodic")]
    cases = []
    for k, (nx, ny, bx, by) in enumerate(shapes):
        U = riemann(k + 21, max(nx, ny))[0][:, :ny, :nx]
        cases.append((oracle2d.Grid.make(nx, ny, y=(0.0, ny / nx), bc_x=bx, bc_y=by), np.ascontiguousarray(U)))
    results = {}
    for key in (dict(seams=2, rows_per_task=0), dict(seams=2, rows_per_task=9, waves_per_block=2),
                dict(seams=2, rows_per_task=64, waves_per_block=1, use_graph=0), dict(seams=0, rows_per_task=16),
                dict(seams=2, rows_per_task=128)):
        with npde_mod.Batch([to_npde_grid(npde_mod, g) for g, _ in cases], scheme, lim, 0.5) as b:
            b.set_tuning(row_groups=0, **key)
            # seam bands keep their seam solves in LDS: at most 64 rows
            assert b.task_layout == ("seams" if key["seams"] and key["rows_per_task"] <= 64 else "strips")
            for k, (_, U) in enumerate(cases):
                b.set_state(k, U)
            b.advance_to(T)
            results[str(key)] = [(b.time(k), b.get_state(k)) for k in range(len(cases))]
    for k, (g, U) in enumerate(cases):
        a = U.copy()
        ro = oracle.run(g, scheme, lim, 0.5, a, T)
        for key, res in results.items():
            (t, steps), got = res[k]
            assert steps == ro["steps"] and t == ro["t"], (key, k)
            assert same(got, a), (key, k)


def test_seams_step_a_state_written_in_place_between_calls(npde_mod, oracle):
    """Seam mode keeps copies of the strips' edge columns between the stage kernels.  A state written in
    place through the device pointer between two calls (which the library cannot see) must still be the
    state the next call steps: the first seam pass of every call gathers from the state itself."""
    import torch
    from npde.families import riemann
    n, nsteps = 128, 6
    g = oracle2d.Grid.make(n, n, bc_x="transmissive", bc_y="periodic")
    UA = np.ascontiguousarray(riemann(31, n)[0])
    UB = np.ascontiguousarray(riemann(32, n)[0])
    with npde_mod.Batch([to_npde_grid(npde_mod, g)], "hllc", "vanleer", 0.5) as b:
        b.set_tuning(row_groups=0, seams=2, steps_per_graph=4)
        assert b.task_layout == "seams"
        b.set_state(0, UA)
        b.advance_steps(nsteps, 0.5)
        view = b.state_view()
        view[0].copy_(torch.from_numpy(UB).to(view.device))
        torch.cuda.synchronize()
        b.advance_steps(nsteps, 0.5)
        got = b.get_state(0)
    a = UB.copy()
    for _ in range(nsteps):
        dt = oracle.compute_dt(g, 0.5, a)[0]
        oracle.step(g, "hllc", "vanleer", a, dt)
    assert same(got, a)


def test_seams_host_driven_steps_match_oracle(npde_mod, oracle):
    """Solver.compute_dt / Solver.step, one host call per step, in seam mode (every call's stage 1
    gathers its seam columns from the state, stage 2 reads the edge arrays): the oracle's dt sequence
    and bits, on a transmissive-x, periodic-y grid with Rusanov + minmod."""
    from npde.families import riemann
    n, ny, nsteps = 192, 80, 5
    g = oracle2d.Grid.make(n, ny, y=(0.0, ny / n), bc_x="transmissive", bc_y="periodic")
    U0 = np.ascontiguousarray(riemann(41, n)[0][:, :ny, :])
    a = U0.copy()
    with npde_mod.Batch([to_npde_grid(npde_mod, g)], "rusanov", "minmod", 0.5) as b:
        b.set_tuning(seams=2, rows_per_task=16)
        assert b.task_layout == "seams"
        b.set_state(0, U0)
        for _ in range(nsteps):
            dt = float(b.compute_dts()[0])
            dto = oracle.compute_dt(g, 0.5, a)[0]
            assert dt == dto
            b.step_all([dt])
            oracle.step(g, "rusanov", "minmod", a, dto)
        assert same(b.get_state(0), a)


def test_tail_retiling_keeps_every_instance_exact(npde_mod, oracle):
    """An uneven batch whose large instances step long after the small ones have finished: the device
    loop rebuilds its task list for the instances still stepping (new band height and layout) in the
    middle of the run.  Every instance's state, frames and frame times equal its own single-instance run,
    and a 256^2 instance equals the oracle."""
    from npde.families import riemann
    sizes = [64, 64, 256, 64, 64, 256, 64, 64]
    ics = [np.ascontiguousarray(riemann(50 + k, n)[0]) for k, n in enumerate(sizes)]
    grids = [npde_mod.make_grid(n, n, bc_x=1, bc_y=k % 2) for k, n in enumerate(sizes)]
    T, NOUT = 0.03, 3

    def run(gs, us, **tuning):
        with npde_mod.Batch(gs, "hllc", "vanleer", 0.5) as b:
            b.set_tuning(steps_per_graph=4, **tuning)
            for k, U in enumerate(us):
                b.set_state(k, U)
            b.run_trajectory(T, NOUT, sizes=(32,))
            return [(b.time(k), b.get_state(k), b.trajectory(k, 32), b.frame_times(k)) for k in range(len(gs))]

    batch = run(grids, ics)
    steps = [r[0][1] for r in batch]
    assert max(steps) >= 3 * min(steps)   # a long tail: the 256^2 instances take several times the steps
    for k in range(len(sizes)):
        (t1, s1), U1, F1, ft1 = run([grids[k]], [ics[k]])[0]
        (t0, s0), U0, F0, ft0 = batch[k]
        assert (t0, s0) == (t1, s1), k
        assert same(U0, U1) and same(F0, F1) and same(ft0, ft1), k
    k = 2
    a = ics[k].copy()
    g = oracle2d.Grid.make(256, 256, bc_x="periodic", bc_y="transmissive")
    ro = oracle.run(g, "hllc", "vanleer", 0.5, a, T, nout=NOUT)
    assert batch[k][0][1] == ro["steps"] and same(batch[k][1], a)


def test_seam_layout_is_the_default_for_wide_instances(npde_mod):
    """The default tuning runs batches of at least 2^18 cells whose widths are multiples of 64 (the
    benchmark batch among them) in seam mode, except 64-wide periodic instances (row groups of one
    wavefront); smaller batches and other widths run in 60-column strips."""
    def layout(nx, bc_x, count, **tuning):
        grids = [npde_mod.make_grid(nx, 64, bc_x=bc_x, bc_y=1)] * count
        with npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as b:
            if tuning:
                b.set_tuning(**tuning)
            return b.task_layout
    assert layout(512, 1, 32) == "seams"
    assert layout(256, 0, 64) == "seams"
    assert layout(256, 1, 64) == "seams"
    assert layout(64, 1, 256) == "row_groups"
    assert layout(64, 0, 256) == "seams"
    assert layout(100, 1, 256) == "strips"
    assert layout(256, 0, 1) == "strips"              # 2^14 cells: too small to fill the GPU
    assert layout(256, 0, 1, seams=2) == "seams"
    assert layout(512, 0, 8) == "seams"               # 2^18 cells


def test_trajectory_invariants_of_check_trajectory(npde_mod):
    """tests/check_trajectory.py:51-74 on a device-loop trajectory (periodic 128^2 random Riemann,
    levels 64/32/16): equally spaced frame times, every coarse level the block average of the finest to
    1e-6, totals equal across levels and constant in time, positive density and pressure."""
    from npde.families import riemann
    with npde_mod.Batch([npde_mod.make_grid(128, 128, bc_x=1, bc_y=1)], "hllc", "vanleer", 0.5) as b:
        b.set_state(0, riemann(11, 128)[0])
        b.run_trajectory(0.2, 10, sizes=(64, 32, 16))
        t = b.frame_times(0)
        L = {m: b.trajectory(0, m).astype(np.float64) for m in (64, 32, 16)}
    assert np.allclose(np.diff(t), 0.02, rtol=0, atol=1e-12)
    fine = L[64]
    for m in (32, 16):
        f = 64 // m
        ba = fine.reshape(11, 4, m, f, m, f).mean(axis=(3, 5))
        assert np.abs(ba - L[m]).max() <= 1e-6 * np.abs(fine).max(), m
        assert np.allclose(L[m].mean(axis=(2, 3)), fine.mean(axis=(2, 3)), rtol=1e-6, atol=1e-6)
    tot = fine.mean(axis=(2, 3))
    assert np.abs(tot - tot[0]).max() <= 1e-6 * np.abs(tot[0]).max()
    rho = fine[:, 0]
    p = 0.4 * (fine[:, 3] - 0.5 * (fine[:, 1] ** 2 + fine[:, 2] ** 2) / rho)
    assert rho.min() > 0 and p.min() > 0

"""Pins the CPU oracle (oracle/euler2d_oracle.cpp) before it is trusted as the GPU's checker.

1. Golden vectors written by the reference's own code (tests/golden/make_golden.py): point-wise
   known answers, whole runs with their dt sequences, a non-physical failure, a CLI trajectory.
2. The reference build itself (oracle/_ref), where present, on inputs the fixtures do not hold.
3. What the mathematics fixes: consistency F(U, U) = F(U), rotation symmetry of the y faces, exact
   x/y transposition of Sod, free-stream preservation, conservation, the exact Sod plateau, the
   published convergence table (32^2 to 256^2, L1 errors and the van Leer L-infinity order), the
   published Kelvin-Helmholtz growth table and step counts, and closed forms of the limiters.
"""
import numpy as np
import pytest

from conftest import run_case_grid

import oracle2d
from conftest import bitwise_equal

SCHEMES = ["rusanov", "hllc"]
LIMS = ["none", "unlimited", "minmod", "vanleer"]


def same(a, b):
    """Bit-for-bit equality (+0 and -0 differ); NaN matches NaN (failed states hold NaNs)"""
    return bitwise_equal(a, b)


# ---------------------------------------------------------------- 1. golden vectors
def test_pointwise_known_answers(oracle, golden):
    k = golden("kat.npz")
    assert same(oracle.to_conserved(k["W"]), k["UL"])
    assert same(oracle.to_conserved(k["W2"]), k["UR"])
    assert same(oracle.to_primitive(k["UL"]), k["prim_UL"])
    assert same(oracle.sound_speed(k["W"]), k["sound_W"])
    assert same(oracle.physical_flux(k["UL"]), k["flux_x_UL"])
    for lim in LIMS:
        assert same(oracle.limiter(lim, k["lim_a"], k["lim_b"]), k[f"lim_{lim}"]), lim
    for s in SCHEMES:
        for d in (0, 1):
            assert same(oracle.face_flux(s, d, k["UL"], k["UR"]), k[f"{s}_{'xy'[d]}"]), (s, d)


RUN_NAMES = ["lax_liu3_64_hllc_vanleer", "lax_liu3_48_rusanov_minmod", "sod_x_64_hllc_unlimited",
             "sod_y_64_rusanov_vanleer", "sine2d_32_hllc_none", "kh_32_rusanov_vanleer", "kh_40_hllc_minmod",
             "riemann_48_hllc_vanleer", "riemann_40_hllc_minmod"]


@pytest.mark.parametrize("name", RUN_NAMES)
def test_whole_runs_bitwise(oracle, golden, name):
    r = golden("runs.npz")
    g = run_case_grid(r[f"{name}/grid"])
    test, scheme, lim = (str(x) for x in r[f"{name}/names"])
    t_end, nout = r[f"{name}/cfg"]
    U = r[f"{name}/U0"].copy()
    out = oracle.run(g, scheme, lim, 0.5, U, float(t_end), nout=int(nout), record_dts=True, max_steps=100000)
    assert out["rc"] == 0
    assert out["steps"] == int(r[f"{name}/steps"])
    assert same(out["dts"], r[f"{name}/dts"])
    assert out["t"] == float(r[f"{name}/t"])
    assert same(out["frame_t"], r[f"{name}/frame_t"])
    assert same(U, r[f"{name}/U"])


def test_builtin_initial_conditions(oracle, golden):
    r = golden("runs.npz")
    for name in RUN_NAMES:
        test = str(r[f"{name}/names"][0])
        if test == "riemann":
            continue
        n = int(r[f"{name}/grid"][0] if test != "sod_y" else r[f"{name}/grid"][1])
        g, _, U0 = oracle.setup(test, n)
        assert same(U0, r[f"{name}/U0"]), name


def test_nonphysical_state_is_reported(oracle, golden):
    r = golden("runs.npz")
    g, _, U0 = oracle.setup("sod_x", 100)
    assert same(U0, r["fail_sod_unlimited/U0"])
    U = U0.copy()
    out = oracle.run(g, "rusanov", "unlimited", 0.5, U, 0.2, nout=1)
    assert out["rc"] == 1
    assert out["bad"] == tuple(int(v) for v in r["fail_sod_unlimited/bad"])
    assert out["steps"] == int(r["fail_sod_unlimited/steps"])
    assert same(U, r["fail_sod_unlimited/U"])


def test_trajectory_block_averages_match_reference_cli(oracle, golden):
    t = golden("traj_ll64.npz")
    g, t_end, U = oracle.setup("lax_liu3", 64)
    out = oracle.run(g, "hllc", "vanleer", 0.5, U, t_end, nout=4, sizes=(64, 32, 16))
    assert same(out["frame_t"], t["t"])
    for m in (64, 32, 16):
        assert out["frames"][m].dtype == np.float32
        assert same(out["frames"][m], t[f"U_{m}"]), m


def test_published_step_count_lax_liu_256(oracle):
    """experiments/results.md:92-94: 682 steps at 256^2, t = 0.3, nout = 10; and results.md:86-88 (README.md:64):
    the final density is symmetric about x = y to ~3e-14 (to one digit: not exact, the x-then-y accumulation
    of the update breaks the mirror symmetry at rounding level)."""
    g, t_end, U = oracle.setup("lax_liu3", 256)
    out = oracle.run(g, "hllc", "vanleer", 0.5, U, t_end, nout=10)
    assert out["steps"] == 682
    assert f"{np.abs(U[0] - U[0].T).max():.0e}" == "3e-14"


# ---------------------------------------------------------------- 2. the reference build itself
@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("lim", LIMS)
def test_against_reference_build_random_riemann(oracle, reference, scheme, lim):
    from npde.families import riemann
    g = oracle2d.Grid.make(36, 28, y=(0.0, 28 / 36), bc_x="periodic", bc_y="transmissive")
    U0, _, _ = riemann(3, 36)
    U0 = np.ascontiguousarray(U0[:, :28, :])
    a, b = U0.copy(), U0.copy()
    ra = oracle.run(g, scheme, lim, 0.5, a, 0.05, nout=2, record_dts=True)
    rb = reference.run(g, scheme, lim, 0.5, b, 0.05, nout=2, record_dts=True)
    assert ra["rc"] == rb["rc"]
    assert ra["steps"] == rb["steps"]
    assert same(ra["dts"], rb["dts"])
    assert same(a, b)


TINY_GRIDS = [(1, 1, "periodic", "periodic"), (1, 5, "transmissive", "periodic"), (5, 1, "periodic", "transmissive"),
              (2, 3, "transmissive", "transmissive"), (3, 2, "periodic", "periodic"), (4, 4, "periodic", "transmissive"),
              (5, 7, "transmissive", "periodic")]


def tiny_state(nx, ny, seed):
    """A positive, non-uniform state on an nx x ny grid (conserved planes)."""
    from npde.solver import to_conserved
    rng = np.random.default_rng(seed)
    rho, u, v, p = (rng.uniform(lo, hi, (ny, nx)) for lo, hi in ((0.2, 2.0), (-0.5, 0.5), (-0.5, 0.5), (0.2, 2.0)))
    return np.ascontiguousarray(to_conserved(rho, u, v, p))


@pytest.mark.parametrize("nx,ny,bx,by", TINY_GRIDS)
def test_grids_narrower_than_the_stencil_against_reference_build(oracle, reference, nx, ny, bx, by):
    """Grids of 1-5 cells per side, where every ghost of the 5-point stencil wraps or clamps onto one or two
    interior cells (grid2d.hpp:43-66 fills x ghosts k = 1, 2 in order, so a 1-cell periodic grid copies
    its only cell twice): the oracle's run equals the reference build's, step by step."""
    g = oracle2d.Grid.make(nx, ny, y=(0.0, ny / nx), bc_x=bx, bc_y=by)
    U0 = tiny_state(nx, ny, 7 * nx + ny)
    a, b = U0.copy(), U0.copy()
    ra = oracle.run(g, "hllc", "vanleer", 0.5, a, 0.05, nout=2, record_dts=True)
    rb = reference.run(g, "hllc", "vanleer", 0.5, b, 0.05, nout=2, record_dts=True)
    assert ra["rc"] == rb["rc"] == 0
    assert ra["steps"] == rb["steps"] > 0
    assert same(ra["dts"], rb["dts"])
    assert same(a, b)


ANISOTROPIC = [(32, 24, (0.0, 1.0), (0.0, 2.25), "periodic", "transmissive"),    # dy = 3 dx
               (40, 40, (-1.0, 1.0), (0.5, 1.0), "transmissive", "periodic")]    # dy = dx / 4


@pytest.mark.parametrize("nx,ny,xr,yr,bx,by", ANISOTROPIC)
def test_anisotropic_cells_against_reference_build(oracle, reference, nx, ny, xr, yr, bx, by):
    """Cells with dy != dx on offset domains: the x and y flux divergences and the CFL dt
    (solver2d.hpp:67-83, 129-180) weigh each direction by its own spacing."""
    from npde.families import riemann
    g = oracle2d.Grid.make(nx, ny, x=xr, y=yr, bc_x=bx, bc_y=by)
    U0 = np.ascontiguousarray(riemann(11, max(nx, ny))[0][:, :ny, :nx])
    a, b = U0.copy(), U0.copy()
    ra = oracle.run(g, "hllc", "vanleer", 0.5, a, 0.04, nout=2, record_dts=True)
    rb = reference.run(g, "hllc", "vanleer", 0.5, b, 0.04, nout=2, record_dts=True)
    assert ra["rc"] == rb["rc"] == 0
    assert ra["steps"] == rb["steps"] > 0
    assert same(ra["dts"], rb["dts"])
    assert same(a, b)


def test_thread_count_does_not_change_bits(oracle, reference):
    """tests/compare_runs.py:11-22's bar across OpenMP thread counts: the oracle with 1, 3 and 8 threads and
    the reference build with 1 thread and with all of them give the same dt sequence and state bits."""
    from npde.families import riemann
    g = oracle2d.Grid.make(48, 40, y=(0.0, 40 / 48), bc_x="periodic", bc_y="transmissive")
    U0 = np.ascontiguousarray(riemann(9, 48)[0][:, :40, :])
    runs = []
    for th in (1, 3, 8):
        a = U0.copy()
        r = oracle.run(g, "hllc", "vanleer", 0.5, a, 0.03, nout=2, record_dts=True, threads=th)
        runs.append((r["steps"], r["dts"], a))
    for th in (1, 0):
        a = U0.copy()
        r = reference.run(g, "hllc", "vanleer", 0.5, a, 0.03, nout=2, record_dts=True, threads=th)
        runs.append((r["steps"], r["dts"], a))
    for steps, dts, a in runs[1:]:
        assert steps == runs[0][0] and same(dts, runs[0][1]) and same(a, runs[0][2])


def test_benchmark_sizes_against_reference_build(oracle, reference):
    """The oracle equals the reference build at the benchmark's own sizes, so the GPU's bitwise checks against
    the oracle at those sizes reach the reference: Lax-Liu 256^2 to t = 0.3 with 10 frames (the reference's CPU
    case, 682 steps), a random-Riemann 512^2 instance of the benchmark batch over its first 0.02 of time, and an
    instance of the held-out Kelvin-Helmholtz family at 256^2 over its first 0.05, and a 256^2 instance of the
    uneven blast batch to its end time 0.2 (1,441 steps) -- every dt, the frame times and the final state bit
    for bit."""
    from npde.families import blast, blast_size, kh, riemann
    cases = []
    g, t_end, U = oracle.setup("lax_liu3", 256)
    cases.append((g, U, 0.0, t_end, 10))
    cases.append((oracle2d.Grid.make(512, 512, bc_x="periodic", bc_y="periodic"), riemann(0, 512)[0], 0.0, 0.02, 2))
    cases.append((oracle2d.Grid.make(256, 256, bc_x="periodic", bc_y="periodic"), kh(0, 256)[0], 0.0, 0.05, 2))
    assert blast_size(5) == 256  # the uneven blast batch's first 256^2 instance, over its whole run to t = 0.2
    cases.append((oracle2d.Grid.make(256, 256, bc_x="periodic", bc_y="periodic"), blast(5)[0], 0.0, 0.2, 2))
    for g, U0, t0, t1, nout in cases:
        a, b = np.ascontiguousarray(U0).copy(), np.ascontiguousarray(U0).copy()
        ra = oracle.run(g, "hllc", "vanleer", 0.5, a, t1, t_start=t0, nout=nout, record_dts=True)
        rb = reference.run(g, "hllc", "vanleer", 0.5, b, t1, t_start=t0, nout=nout, record_dts=True)
        assert ra["rc"] == rb["rc"] == 0 and ra["steps"] == rb["steps"] > 0, g.nx
        assert same(ra["dts"], rb["dts"]) and same(ra["frame_t"], rb["frame_t"]), g.nx
        assert same(a, b), g.nx


def test_random_configurations_against_reference_build(oracle, reference):
    """The 64 seeded configurations of the GPU test (grid, boundaries, scheme, limiter; launch shapes do
    not apply on the CPU): the oracle's run equals the reference build's -- dt sequence, step count and
    bits, or the same first non-physical cell."""
    from test_gpu_parity import random_case
    for seed in range(64):
        g, U, scheme, lim, _ = random_case(np.random.default_rng(1000 + seed))
        a, b = U.copy(), U.copy()
        ra = oracle.run(g, scheme, lim, 0.5, a, 0.03, nout=3, record_dts=True)
        rb = reference.run(g, scheme, lim, 0.5, b, 0.03, nout=3, record_dts=True)
        assert ra["rc"] == rb["rc"] and ra["steps"] == rb["steps"], seed
        assert ra["bad"] == rb["bad"], seed
        assert same(ra["dts"], rb["dts"]), seed
        if ra["bad"] is None:
            assert same(a, b), seed


def test_against_reference_build_pointwise_random(oracle, reference):
    rng = np.random.default_rng(5)
    n = 20000
    W = np.stack([rng.uniform(1e-3, 10, n), rng.normal(0, 3, n), rng.normal(0, 3, n), rng.uniform(1e-3, 10, n)], 1)
    W2 = np.stack([rng.uniform(1e-3, 10, n), rng.normal(0, 3, n), rng.normal(0, 3, n), rng.uniform(1e-3, 10, n)], 1)
    UL, UR = reference.to_conserved(W), reference.to_conserved(W2)
    for s in SCHEMES:
        for d in (0, 1):
            assert same(oracle.face_flux(s, d, UL, UR), reference.face_flux(s, d, UL, UR))


# ---------------------------------------------------------------- 3. what the mathematics fixes
def test_limiter_closed_forms(oracle):
    a = np.array([1.0, 2.0, -1.0, 1.0, 0.0, 3.0, -2.0, 0.5])
    b = np.array([2.0, 1.0, 1.0, -1.0, 1.0, 3.0, -4.0, 0.5])
    assert same(oracle.limiter("none", a, b), np.zeros(8))
    assert same(oracle.limiter("unlimited", a, b), 0.5 * (a + b))
    assert same(oracle.limiter("minmod", a, b), [1.0, 1.0, 0.0, 0.0, 0.0, 3.0, -2.0, 0.5])
    # van Leer: harmonic mean 2ab/(a+b) for same signs, else 0 (TVD); symmetric in (a, b)
    vl = oracle.limiter("vanleer", a, b)
    assert np.allclose(vl, [4 / 3, 4 / 3, 0.0, 0.0, 0.0, 3.0, -8 / 3, 0.5], rtol=0, atol=1e-15)
    assert same(oracle.limiter("vanleer", b, a), vl)


def test_flux_consistency_and_rotation(oracle):
    """test_riemann2d.cpp:42-47 (HLLC(U,U) = F(U)) and :33-39 (y face = x face with momenta swapped)."""
    W = np.array([[1.3, 0.4, -0.7, 2.1], [0.2, 3.0, 0.1, 0.05], [2.0, -1.5, 0.7, 4.0]])
    U = oracle.to_conserved(W)
    F = oracle.physical_flux(U)
    for s in SCHEMES:
        got = oracle.face_flux(s, 0, U, U)
        assert np.allclose(got, F, rtol=1e-13, atol=1e-13), s
    sw = lambda X: X[:, [0, 2, 1, 3]]
    rng = np.random.default_rng(2)
    Wl = np.stack([rng.uniform(0.1, 2, 50), rng.normal(size=50), rng.normal(size=50), rng.uniform(0.1, 2, 50)], 1)
    Wr = np.stack([rng.uniform(0.1, 2, 50), rng.normal(size=50), rng.normal(size=50), rng.uniform(0.1, 2, 50)], 1)
    UL, UR = oracle.to_conserved(Wl), oracle.to_conserved(Wr)
    for s in SCHEMES:
        assert same(oracle.face_flux(s, 1, sw(UL), sw(UR)), sw(oracle.face_flux(s, 0, UL, UR)))


def test_sod_transpose_symmetry_conservation_and_plateau(oracle):
    """test_solver2d.cpp:61-108 restated on the oracle."""
    gx, t_end, Ux = oracle.setup("sod_x", 200)
    gy, _, Uy = oracle.setup("sod_y", 200)
    mass0 = Ux[0].sum() * (1 / 200) * (4 / 200)
    en0 = Ux[3].sum() * (1 / 200) * (4 / 200)
    rx = oracle.run(gx, "hllc", "vanleer", 0.5, Ux, t_end)
    ry = oracle.run(gy, "hllc", "vanleer", 0.5, Uy, t_end)
    assert rx["steps"] == ry["steps"]
    assert (Ux[:, 1:, :] == Ux[:, :1, :]).all()            # rows identical
    assert (Ux[2] == 0.0).all()                              # no y-momentum created
    T = np.transpose(Uy, (0, 2, 1))[[0, 2, 1, 3]]            # y-run transposed, momenta swapped
    assert np.abs(Ux - T).max() == 0.0
    assert abs(Ux[0].sum() * (1 / 200) * (4 / 200) - mass0) < 1e-13
    assert abs(Ux[3].sum() * (1 / 200) * (4 / 200) - en0) < 1e-13
    i = int(0.78 * 200)
    rho, mx, _, E = Ux[:, 0, i]
    u = mx / rho
    p = 0.4 * (E - 0.5 * rho * u * u)
    assert abs(rho - 0.26557) < 1e-3 and abs(p - 0.30313) < 1e-3 and abs(u - 0.92745) < 1e-3


def test_free_stream_preserved(oracle):
    """test_solver2d.cpp:36-53: a uniform moving state stays uniform on a periodic 16 x 12 grid."""
    g = oracle2d.Grid.make(16, 12, bc_x="periodic", bc_y="periodic")
    U0 = oracle.to_conserved(np.array([[1.3, 0.7, -0.4, 2.0]]))[0]
    U = np.ascontiguousarray(np.broadcast_to(U0[:, None, None], (4, 12, 16)))
    oracle.run(g, "hllc", "vanleer", 0.5, U, 0.5)
    assert np.abs(U - U0[:, None, None]).max() < 1e-13


# experiments/results.md:71-80: L1 density error of the diagonal sine wave at t = 1 (printed to 4 digits)
# and the L-infinity order of van Leer between successive sizes (printed to 2 decimals)
SINE2D_L1 = {("vanleer", 32): 1.088e-2, ("vanleer", 64): 3.203e-3, ("vanleer", 128): 8.531e-4,
             ("vanleer", 256): 2.130e-4, ("none", 32): 9.024e-2, ("none", 64): 5.862e-2,
             ("none", 128): 3.379e-2, ("none", 256): 1.820e-2}


@pytest.fixture(scope="module")
def sine2d_errors(oracle):
    """(L1, L-infinity) density errors of every run of the published table."""
    errs = {}
    for (lim, n) in SINE2D_L1:
        g, t_end, U = oracle.setup("sine2d", n)
        out = oracle.run(g, "hllc", lim, 0.5, U, t_end, nout=1)
        x = (np.arange(n) + 0.5) / n
        X, Y = np.meshgrid(x, x)
        exact = 1.0 + 0.2 * np.sin(2 * np.pi * (X + Y - 2 * out["t"]))
        e = np.abs(U[0] - exact)
        errs[(lim, n)] = (e.sum() * (1 / n) * (1 / n), e.max())
    return errs


@pytest.mark.parametrize("lim,n", list(SINE2D_L1))
def test_published_convergence_table(sine2d_errors, lim, n):
    """The diagonal-advection table of experiments/results.md:71-80, HLLC, SSP-RK2, CFL 0.5, 32^2 to 256^2."""
    assert float(f"{sine2d_errors[(lim, n)][0]:.3e}") == SINE2D_L1[(lim, n)]


def test_published_linf_order_of_van_leer(sine2d_errors):
    """results.md:75-80: the L-infinity order of van Leer stays at 1.41 from 64^2 to 256^2 (clipping at the
    smooth extrema), while its L1 order approaches 2."""
    for n in (64, 128, 256):
        order = np.log2(sine2d_errors[("vanleer", n // 2)][1] / sine2d_errors[("vanleer", n)][1])
        assert round(float(order), 2) == 1.41, n
    assert round(float(np.log2(sine2d_errors[("vanleer", 128)][0] / sine2d_errors[("vanleer", 256)][0])), 2) == 2.00


def kh_mode_amplitude(U, n):
    """The seeded mode's amplitude by the diagnostic of experiments/kh_growth.py:17-27 (McNally, Lyra &
    Passy 2012, eqs. 4-7): v projected on sin(4 pi x) and cos(4 pi x), weighted by exp(-4 pi d) with d the
    distance to the nearer shear layer."""
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    v = U[2].astype(np.float64) / U[0].astype(np.float64)
    dist = np.where(Y < 0.5, np.abs(Y - 0.25), np.abs((1.0 - Y) - 0.25))
    w = np.exp(-4 * np.pi * dist)
    s, c, d = np.sum(v * np.sin(4 * np.pi * X) * w), np.sum(v * np.cos(4 * np.pi * X) * w), np.sum(w)
    return 2 * np.sqrt((s / d) ** 2 + (c / d) ** 2)


@pytest.mark.parametrize("n,rate,final", [(16, None, None), (32, 1.20, None), (64, 2.41, 0.193), (128, 2.71, 0.225),
                                           (256, 2.77, 0.228)])
def test_published_kh_growth_table(oracle, n, rate, final):
    """experiments/results.md:96-102 and README.md:66-71: the built-in Kelvin-Helmholtz test (main2d.cpp:98-129),
    HLLC + van Leer to t = 2, growth rate of the seeded mode fitted over 0.6 <= t <= 1.2 (kh_growth.py:37-39)
    and its final amplitude, as printed (2 and 3 digits), and at 16^2 a mode that decays.  Neither source
    gives the snapshot count; 40 frames reproduce all eight figures (20 frames give 1.21 at 32^2; 10 frames
    2.40, 2.69 and 2.75 from 64^2 up: the fit then sees 4 points).  A long-horizon pin of the shear-layer
    initial condition and of the solver's numerical diffusion, which sets how much the coarse grids
    under-predict the growth."""
    g, t_end, U = oracle.setup("kh", n)
    nout = 40
    out = oracle.run(g, "hllc", "vanleer", 0.5, U, t_end, nout=nout, sizes=(n,))
    t = out["frame_t"]
    amp = np.array([kh_mode_amplitude(f, n) for f in out["frames"][n]])
    assert t[-1] == 2.0 and amp[0] == pytest.approx(0.01, rel=1e-3)  # the 0.01 sin(4 pi x) seed
    lin = (t >= 0.6) & (t <= 1.2)
    fit = float(np.polyfit(t[lin], np.log(amp[lin]), 1)[0])
    if rate is None:  # "decays"
        assert fit < 0.0 and amp[-1] < 0.1 * amp[0]
    else:
        assert round(fit, 2) == rate
    if final is not None:
        assert round(float(amp[-1]), 3) == final


def test_lax_liu_near_symmetry(oracle):
    """README.md:64: Lax-Liu 3 is symmetric about x = y to ~3e-14 (x-then-y accumulation breaks exactness)."""
    g, t_end, U = oracle.setup("lax_liu3", 64)
    oracle.run(g, "hllc", "vanleer", 0.5, U, t_end)
    d = np.abs(U[0] - U[0].T).max()
    assert 0.0 < d < 1e-12

"""The euler2d-compatible driver (npde-euler2d): flags, messages, exit codes and HDF5 outputs against the
reference CLI's golden outputs (tests/golden/cli.json, traj_ll64.npz)."""
import json
import os
import re
import shutil
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "neural-pde-surrogate_amd" / "lib" / "npde-euler2d"
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
GOLDEN = Path(__file__).resolve().parent / "golden"


def run(*args, cwd=None):
    if not EXE.exists():
        pytest.skip("driver not built")
    return subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=600)


def h5(path, name, dtype):
    if not H5DUMP:
        pytest.skip("h5dump not available")
    with tempfile.NamedTemporaryFile(suffix=".bin") as tmp:
        subprocess.run([H5DUMP, "-b", "LE", "-d", name, "-o", tmp.name, str(path)], check=True,
                       stdout=subprocess.DEVNULL)
        return np.fromfile(tmp.name, dtype=dtype)


@pytest.mark.parametrize("args,msg", [
    (["--bogus", "1"], "Unknown option --bogus"),
    (["--n"], "Expected '--key value', got '--n'"),
    (["-n", "5"], "Expected '--key value', got '-n'"),
    (["--n", "64", "extra"], "Expected '--key value', got 'extra'"),
    (["--test", "nope"], "Unknown test 'nope' (use sod_x, sod_y, lax_liu3, sine2d or kh)"),
    (["--scheme", "roe"], "Unknown scheme 'roe' (use rusanov or hllc)"),
    (["--limiter", "superbee"], "Unknown limiter 'superbee'"),
    (["--n", "0"], "n, nout and cfl must be positive"),
    (["--cfl", "-1"], "n, nout and cfl must be positive"),
    (["--tend", "-2"], "tend must be positive"),
])
def test_argument_errors_match_reference(args, msg):
    """main2d.cpp:24-43, 138-179: same messages, exit code 1, before any device work."""
    p = run(*args)
    assert p.returncode == 1
    assert p.stderr.strip() == msg


@pytest.mark.gpu
def test_trajectory_file_matches_reference_cli(npde_mod, tmp_path):
    t = np.load(GOLDEN / "traj_ll64.npz")
    ref_out = json.loads((GOLDEN / "cli.json").read_text())["ll64_traj_stdout"]
    out = tmp_path / "ll64.h5"
    p = run("--test", "lax_liu3", "--n", "64", "--nout", "4", "--traj", out, "--save", "64,32,16")
    assert p.returncode == 0, p.stderr
    steps = int(re.search(r"Done: (\d+) steps", p.stdout).group(1))
    assert steps == int(re.search(r"Done: (\d+) steps", ref_out).group(1))
    assert np.array_equal(h5(out, "t", "<f8"), t["t"])
    for m in (64, 32, 16):
        assert np.array_equal(h5(out, f"U_{m}", "<f4").reshape(5, 4, m, m), t[f"U_{m}"]), m


@pytest.mark.gpu
def test_profile_table_splits_the_run_by_kernel(npde_mod, tmp_path):
    """--profile yes prints the reference's phase table (solver2d.hpp:43-51) with one row per kernel kind
    (launch counts from events around every launch of the graph-free profiled run); the profiled run's
    trajectory is bitwise the unprofiled one."""
    t = np.load(GOLDEN / "traj_ll64.npz")
    out = tmp_path / "ll64.h5"
    p = run("--test", "lax_liu3", "--n", "64", "--nout", "4", "--traj", out, "--save", "64,32,16",
            "--profile", "yes")
    assert p.returncode == 0, p.stderr
    steps = int(re.search(r"Done: (\d+) steps", p.stdout).group(1))
    rows = {m.group(1).strip(): (float(m.group(2)), int(m.group(3)) if m.group(3) else None)
            for m in re.finditer(r"^(\S.*?)\s+([0-9.]+)\s+[0-9.]+%\s+[0-9.]+(?:\s+(\d+))?$", p.stdout, re.M)}
    assert rows["stage 1 (slopes, fluxes, U1)"][1] == rows["stage 2 (+ U^n+1, wave speeds)"][1] >= steps
    # every step is completed by a k_finish or, in launches with frames, by the next frame kernel
    assert (rows["dt reduction + run state"][1] + rows["frames (block averages)"][1]
            > rows["stage 2 (+ U^n+1, wave speeds)"][1])
    assert rows["frames (block averages)"][1] >= 5
    assert rows["stage 1 (slopes, fluxes, U1)"][0] > 0 and rows["stage 2 (+ U^n+1, wave speeds)"][0] > 0
    for k in ("launch gaps (device loop)", "output (copy + HDF5)", "host (setup, sync)", "total"):
        assert k in rows
    assert np.array_equal(h5(out, "t", "<f8"), t["t"])
    for m in (64, 32, 16):
        assert np.array_equal(h5(out, f"U_{m}", "<f4").reshape(5, 4, m, m), t[f"U_{m}"]), m


@pytest.mark.gpu
def test_profiled_batch_matches_the_unprofiled_run(npde_mod):
    """npde_set_profiling: a profiled device-resident run (kernels launched one at a time between events) gives
    the same bits, step counts and frames as the graph-launched run; the profile counts one stage-1 and one
    stage-2 launch per loop iteration and at least one per step, and switching it on again clears it."""
    from npde.families import riemann
    n, B = 128, 3
    ics = [riemann(40 + k, n)[0] for k in range(B)]
    grid = npde_mod.make_grid(n, n, bc_x=1, bc_y=1)
    res = []
    for prof in (False, True):
        with npde_mod.Batch([grid] * B, "hllc", "vanleer", 0.5) as b:
            b.set_profiling(prof)
            for k, U in enumerate(ics):
                b.set_state(k, U)
            b.run_trajectory(0.05, 5, sizes=(64,))
            res.append(([b.get_state(k) for k in range(B)], [b.trajectory(k, 64) for k in range(B)], b.times()[1]))
            pr = b.profile()
            if prof:
                assert pr["stage1"][1] == pr["stage2"][1] >= res[-1][2].max()
                assert pr["finish"][1] + pr["frame"][1] >= pr["stage2"][1]
                assert pr["frame"][1] >= 6 and pr["wavespeed"][1] >= 1
                assert pr["stage1"][0] > 0 and pr["stage2"][0] > 0 and pr["network"] == (0.0, 0)
                b.set_profiling(True)
                assert all(v == (0.0, 0) for v in b.profile().values())
            else:
                assert all(v == (0.0, 0) for v in pr.values())
    (sa, fa, na), (sb, fb, nb) = res
    assert (na == nb).all()
    for k in range(B):
        assert np.array_equal(sa[k].view(np.uint64), sb[k].view(np.uint64))
        assert np.array_equal(fa[k], fb[k])


def layout(path):
    """`h5dump -p -A` of a file (names, types, shapes, storage, filters, attribute values; no dataset
    data) with its path replaced by its base name, as tests/golden/make_golden.py records it."""
    if not H5DUMP:
        pytest.skip("h5dump not available")
    out = subprocess.run([H5DUMP, "-p", "-A", str(path)], capture_output=True, text=True, check=True).stdout
    return out.replace(str(path), Path(path).name)


def same_text(got, want):
    import difflib
    return got == want, "\n".join(difflib.unified_diff(want.splitlines(), got.splitlines(), "reference", "npde",
                                                        lineterm="", n=2))


@pytest.mark.gpu
def test_hdf5_layouts_match_reference_writer(npde_mod, tmp_path):
    """A snapshot (io_hdf5.hpp:56-92) and a trajectory file (trajectory.hpp:53-91) carry the same
    datasets, attributes (names, types, values), storage layouts and filters as the reference CLI's
    (tests/golden/h5_layout.json), and the snapshot's datasets are the reference's bit for bit."""
    gold = json.loads((GOLDEN / "h5_layout.json").read_text())
    p = run("--test", "sine2d", "--n", "32", "--nout", "2", "--out", tmp_path)
    assert p.returncode == 0, p.stderr
    snap = tmp_path / "sine2d_hllc_vanleer_N32" / "snap_0001.h5"
    ok, diff = same_text(layout(snap), gold["snap_sine32"])
    assert ok, diff
    ref = np.load(GOLDEN / "snap_sine32.npz")
    for q in ("rho", "u", "v", "p", "x", "y"):
        assert np.array_equal(h5(snap, q, "<f8"), ref[q]), q
    traj = tmp_path / "ll32.h5"
    p = run("--test", "lax_liu3", "--n", "32", "--nout", "2", "--traj", traj, "--save", "32,16")
    assert p.returncode == 0, p.stderr
    ok, diff = same_text(layout(traj), gold["traj_ll32"])
    assert ok, diff


@pytest.mark.gpu
def test_snapshot_mode_step_count_and_restart(npde_mod, oracle, tmp_path):
    """682 steps at 256^2 / nout 10 (experiments/results.md:94); snapshots are valid initial conditions
    (io_hdf5.hpp:171-175): restarting from one equals the oracle run from the same primitives."""
    p = run("--test", "lax_liu3", "--n", "256", "--nout", "10", "--out", tmp_path)
    assert p.returncode == 0, p.stderr
    assert "Done: 682 steps" in p.stdout
    snaps = sorted((tmp_path / "lax_liu3_hllc_vanleer_N256").glob("snap_*.h5"))
    assert len(snaps) == 11
    p = run("--test", "lax_liu3", "--n", "64", "--nout", "2", "--out", tmp_path)
    assert p.returncode == 0
    snap = tmp_path / "lax_liu3_hllc_vanleer_N64" / "snap_0001.h5"
    rho, u, v, pr = (h5(snap, q, "<f8").reshape(64, 64) for q in ("rho", "u", "v", "p"))
    p2 = run("--init", snap, "--tend", "0.05", "--nout", "1", "--out", tmp_path / "restart")
    assert p2.returncode == 0, p2.stderr
    got = {q: h5(tmp_path / "restart" / "lax_liu3_hllc_vanleer_N64" / "snap_0001.h5", q, "<f8").reshape(64, 64)
           for q in ("rho", "u", "v", "p")}
    import oracle2d
    g = oracle2d.Grid.make(64, 64)
    W = np.stack([rho, u, v, pr], -1).reshape(-1, 4)
    U = np.ascontiguousarray(np.moveaxis(oracle.to_conserved(W).reshape(64, 64, 4), -1, 0))
    oracle.run(g, "hllc", "vanleer", 0.5, U, 0.05, nout=1)
    Wo = oracle.to_primitive(np.moveaxis(U, 0, -1).reshape(-1, 4)).reshape(64, 64, 4)
    for k, q in enumerate(("rho", "u", "v", "p")):
        assert np.array_equal(got[q], Wo[..., k]), q


@pytest.mark.gpu
def test_nonphysical_state_exits_like_reference(npde_mod, tmp_path):
    ref = json.loads((GOLDEN / "cli.json").read_text())
    p = run("--test", "sod_x", "--n", "100", "--limiter", "unlimited", "--scheme", "rusanov", "--traj",
            tmp_path / "f.h5")
    assert p.returncode == ref["sod_unlimited_exit"] == 1
    cell = re.search(r"in cell \((\d+), (\d+)\)", p.stderr).groups()
    assert cell == re.search(r"in cell \((\d+), (\d+)\)", ref["sod_unlimited_stderr"]).groups()
    assert re.fullmatch(r"Non-physical state in cell \(\d+, \d+\): rho = -?nan, p = -?nan\n", p.stderr)

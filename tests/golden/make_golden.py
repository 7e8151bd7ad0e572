"""Write the golden vectors under tests/golden/ from the reference solver's own code.

Runs in the container that holds the reference sources: it builds oracle/_ref (the reference's
headers and driver compiled from solver/src by oracle/Makefile) and records
  kat.npz          point-wise known answers: to_primitive / to_conserved / sound_speed / flux_x,
                   the four limiters, Rusanov and HLLC on x and y faces (state2d.hpp, limiters.hpp,
                   riemann2d.hpp), on random and hand-picked edge-case inputs
  runs.npz         whole runs of Solver::step with the driver's landing loop (main2d.cpp:224-246):
                   initial state, final state, every dt, step count; plus a non-physical failure
  traj_ll64.npz    block-averaged float32 trajectory written by the reference euler2d CLI
                   (--traj --save 64,32,16, trajectory.hpp) for Lax-Liu 64^2, read back with h5dump
  cli.json         step counts and "Done:" lines of the reference CLI (Lax-Liu 256^2, nout 10 -> 682)
  h5_layout.json   `h5dump -p -A` of a snapshot (sine2d 32^2, nout 2, snap_0001.h5) and of a trajectory
                   file (Lax-Liu 32^2, nout 2, --save 32,16) written by the reference CLI: dataset and
                   attribute names, types, shapes, storage and filters, attribute values
  snap_sine32.npz  the datasets of that snapshot (rho, u, v, p, x, y)

Only data is written: inputs and outputs.  Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "neural-pde-surrogate_amd"))
import oracle2d  # noqa: E402

OUT = Path(__file__).resolve().parent
H5DUMP = "/opt/conda/bin/h5dump"


def kat(ref):
    rng = np.random.default_rng(20240611)
    n = 600
    W = np.stack([rng.uniform(0.05, 3.0, n), rng.uniform(-2.5, 2.5, n), rng.uniform(-2.5, 2.5, n),
                  rng.uniform(0.05, 3.0, n)], axis=1)
    W2 = np.stack([rng.uniform(0.05, 3.0, n), rng.uniform(-2.5, 2.5, n), rng.uniform(-2.5, 2.5, n),
                   rng.uniform(0.05, 3.0, n)], axis=1)
    # edge cases: equal states, zero velocities, strong supersonic flows both ways, Sod, colliding
    edge = np.array([[1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1], [1.0, 0.75, 0.0, 1.0], [1.0, 2.0, 0.3, 0.4],
                     [1.0, -2.0, -0.3, 0.4], [1.3, 0.4, -0.7, 2.1], [0.2, 8.0, 1.0, 0.05], [0.2, -8.0, 1.0, 0.05],
                     [2.0, 0.0, 5.0, 0.1], [0.5323, 1.206, 0.0, 0.3]])
    edge2 = np.array([[0.125, 0.0, 0.0, 0.1], [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1], [1.0, -2.0, -0.3, 0.4],
                      [1.0, 2.0, 0.3, 0.4], [1.3, 0.4, -0.7, 2.1], [0.3, 7.0, -1.0, 0.06], [0.3, -7.0, -1.0, 0.06],
                      [2.0, 0.0, -5.0, 0.1], [0.138, 1.206, 1.206, 0.029]])
    W = np.concatenate([W, edge])
    W2 = np.concatenate([W2, edge2])
    UL = ref.to_conserved(W)
    UR = ref.to_conserved(W2)
    a = np.concatenate([rng.normal(size=400), [0.0, 1.0, -1.0, 2.0, 0.5, 0.0, 3.0, -3.0]])
    b = np.concatenate([rng.normal(size=400), [1.0, 0.0, 1.0, -2.0, 0.5, 0.0, -3.0, 3.0]])
    out = {"W": W, "W2": W2, "UL": UL, "UR": UR,
           "prim_UL": ref.to_primitive(UL), "sound_W": ref.sound_speed(W), "flux_x_UL": ref.physical_flux(UL),
           "lim_a": a, "lim_b": b}
    for k in ("none", "unlimited", "minmod", "vanleer"):
        out[f"lim_{k}"] = ref.limiter(k, a, b)
    for s in ("rusanov", "hllc"):
        for d in (0, 1):
            out[f"{s}_{'xy'[d]}"] = ref.face_flux(s, d, UL, UR)
    np.savez_compressed(OUT / "kat.npz", **out)


def riemann_ic(n, index):
    from npde.families import riemann  # numpy-only module of the product package
    U, _, _ = riemann(index, n)
    return U


RUNS = [
    # name, (test or 'riemann'), n, scheme, limiter, t_end, nout
    ("lax_liu3_64_hllc_vanleer", "lax_liu3", 64, "hllc", "vanleer", 0.3, 3),
    ("lax_liu3_48_rusanov_minmod", "lax_liu3", 48, "rusanov", "minmod", 0.15, 2),
    ("sod_x_64_hllc_unlimited", "sod_x", 64, "hllc", "unlimited", 0.2, 1),
    ("sod_y_64_rusanov_vanleer", "sod_y", 64, "rusanov", "vanleer", 0.2, 1),
    ("sine2d_32_hllc_none", "sine2d", 32, "hllc", "none", 0.25, 2),
    ("kh_32_rusanov_vanleer", "kh", 32, "rusanov", "vanleer", 0.3, 1),
    ("kh_40_hllc_minmod", "kh", 40, "hllc", "minmod", 0.2, 1),
    ("riemann_48_hllc_vanleer", "riemann", 48, "hllc", "vanleer", 0.1, 2),
    ("riemann_40_hllc_minmod", "riemann", 40, "hllc", "minmod", 0.08, 1),
]


def runs(orc, ref):
    out = {}
    for name, test, n, scheme, lim, t_end, nout in RUNS:
        if test == "riemann":
            g = oracle2d.Grid.make(n, n, bc_x="periodic", bc_y="periodic")
            U0 = riemann_ic(n, 7)
        else:
            g, _, U0 = orc.setup(test, n)
        U = U0.copy()
        r = ref.run(g, scheme, lim, 0.5, U, t_end, nout=nout, record_dts=True, max_steps=200000)
        assert r["rc"] == 0, name
        out[f"{name}/grid"] = np.array([g.nx, g.ny, g.x_min, g.x_max, g.y_min, g.y_max, g.bc_x, g.bc_y])
        out[f"{name}/U0"] = U0
        out[f"{name}/U"] = U
        out[f"{name}/dts"] = r["dts"]
        out[f"{name}/steps"] = np.array(r["steps"])
        out[f"{name}/t"] = np.array(r["t"])
        out[f"{name}/frame_t"] = r["frame_t"]
        out[f"{name}/cfg"] = np.array([t_end, nout], dtype=np.float64)
        out[f"{name}/names"] = np.array([test, scheme, lim])
    # failure: unlimited slopes on Sod overshoot to negative pressure (experiments/results.md:40-46)
    g, _, U0 = orc.setup("sod_x", 100)
    U = U0.copy()
    r = ref.run(g, "rusanov", "unlimited", 0.5, U, 0.2, nout=1, max_steps=200000)
    assert r["rc"] == 1
    out["fail_sod_unlimited/U0"] = U0
    out["fail_sod_unlimited/bad"] = np.array(r["bad"])
    out["fail_sod_unlimited/steps"] = np.array(r["steps"])
    out["fail_sod_unlimited/U"] = U
    np.savez_compressed(OUT / "runs.npz", **out)


def h5_dataset(path, name, dtype):
    with tempfile.NamedTemporaryFile(suffix=".bin") as tmp:
        subprocess.run([H5DUMP, "-b", "LE", "-d", name, "-o", tmp.name, str(path)], check=True,
                       stdout=subprocess.DEVNULL)
        return np.fromfile(tmp.name, dtype=dtype)


def cli(ref_cli):
    info = {}
    with tempfile.TemporaryDirectory() as d:
        traj = Path(d) / "ll64.h5"
        p = subprocess.run([str(ref_cli), "--test", "lax_liu3", "--n", "64", "--nout", "4", "--traj", str(traj),
                            "--save", "64,32,16"], capture_output=True, text=True, check=True)
        info["ll64_traj_stdout"] = p.stdout
        frames = {m: h5_dataset(traj, f"U_{m}", "<f4").reshape(5, 4, m, m) for m in (64, 32, 16)}
        t = h5_dataset(traj, "t", "<f8")
        np.savez_compressed(OUT / "traj_ll64.npz", U_64=frames[64], U_32=frames[32], U_16=frames[16], t=t)
        p = subprocess.run([str(ref_cli), "--test", "lax_liu3", "--n", "256", "--nout", "10", "--out", d],
                           capture_output=True, text=True, check=True)
        info["ll256_nout10_stdout"] = p.stdout
        p = subprocess.run([str(ref_cli), "--test", "sod_x", "--n", "100", "--limiter", "unlimited", "--scheme",
                            "rusanov", "--traj", str(Path(d) / "f.h5")], capture_output=True, text=True)
        info["sod_unlimited_exit"] = p.returncode
        info["sod_unlimited_stderr"] = p.stderr
    (OUT / "cli.json").write_text(json.dumps(info, indent=1) + "\n")


def h5_layout_text(path):
    """`h5dump -p -A` of a file with its path replaced by its base name."""
    out = subprocess.run([H5DUMP, "-p", "-A", str(path)], capture_output=True, text=True, check=True).stdout
    return out.replace(str(path), Path(path).name)


def h5_layouts(ref_cli):
    info = {}
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([str(ref_cli), "--test", "sine2d", "--n", "32", "--nout", "2", "--out", d],
                       capture_output=True, text=True, check=True)
        snap = Path(d) / "sine2d_hllc_vanleer_N32" / "snap_0001.h5"
        info["snap_sine32"] = h5_layout_text(snap)
        np.savez_compressed(OUT / "snap_sine32.npz",
                            **{q: h5_dataset(snap, q, "<f8") for q in ("rho", "u", "v", "p", "x", "y")})
        traj = Path(d) / "ll32.h5"
        subprocess.run([str(ref_cli), "--test", "lax_liu3", "--n", "32", "--nout", "2", "--traj", str(traj),
                        "--save", "32,16"], capture_output=True, text=True, check=True)
        info["traj_ll32"] = h5_layout_text(traj)
    (OUT / "h5_layout.json").write_text(json.dumps(info, indent=1) + "\n")


def main():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "-j8"], check=True)
    ref = oracle2d.Reference()
    orc = oracle2d.Oracle()
    kat(ref)
    runs(orc, ref)
    cli(oracle2d.REF_EULER2D)
    h5_layouts(oracle2d.REF_EULER2D)
    for f in sorted(OUT.glob("*.npz")) + [OUT / "cli.json", OUT / "h5_layout.json"]:
        print(f"{f.name}: {f.stat().st_size / 1024:.1f} KiB")


if __name__ == "__main__":
    main()

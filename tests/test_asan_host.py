"""Host code under sanitizers, on the CPU: the dataset job's HDF5 packer (libnpde_io.so) and the oracle, each
built here by g++ with the sanitizers and driven through its Python callers in a child process."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
IO_LIB = ROOT / "neural-pde-surrogate_amd" / "lib" / "libnpde_io.so"


def preload(*runtimes):
    """LD_PRELOAD for the child: the sanitizer runtimes first (they must load before anything else), then
    whatever the environment already preloads."""
    return ":".join([*runtimes, *filter(None, [os.environ.get("LD_PRELOAD", "")])])


# The dataset job's host-side writes, run in a child process against a build of the writer instrumented by
# AddressSanitizer: packed files of several shapes, converted (non-contiguous, other dtypes) inputs, rows
# with and without a shared frame spacing, and the whole pack step (files, statistics, manifest) on
# synthetic trajectories, each input array released as soon as its call returns.
ASAN_SCRIPT = r"""
import gc, sys, tempfile
from pathlib import Path
import numpy as np
sys.path.insert(0, sys.argv[1])
import npde.dataset as D
rng = np.random.default_rng(0)
with tempfile.TemporaryDirectory() as d:
    for it, (N, F, n) in enumerate([(3, 5, 8), (1, 51, 64), (40, 51, 16), (7, 2, 4)]):
        U = rng.normal(size=(N, F, 4, n, n)).astype(np.float32)
        t = np.tile(np.linspace(0.0, 0.5, F), (N, 1)) + (0.0 if it % 2 else rng.uniform(size=(N, 1)))
        fams = [("riemann", "blast", "kh")[k % 3] * (1 + k % 4) for k in range(N)]
        D.write_packed(Path(d) / f"x{it}.h5", U, t, fams, list(range(N)), "train" if it % 2 else "ood",
                       D.frame_spacing(t))
        del U, t
        gc.collect()
    U = rng.normal(size=(2, 3, 4, 8, 16))[..., ::2]
    D.write_packed(Path(d) / "nc.h5", U, np.zeros((2, 3)), ["riemann", "blast"], np.array([5, 6], np.int64), "val",
                   float("nan"))
    trajs = {}
    for k in range(12):
        fam = "riemann" if k < 8 else "kh"
        frames = {m: np.abs(rng.normal(size=(51, 4, m, m))).astype(np.float32) + 0.5 for m in (16, 8)}
        trajs[(fam, k)] = {"frames": frames, "t": np.linspace(0.0, 0.5, 51) + (0.7 if fam == "kh" else 0.0),
                           "n": 64, "steps": 100 + k}
    D.pack(trajs, (16, 8), Path(d), log=lambda s: None)
    assert len(list(Path(d).iterdir())) == 12
print("PACKER RUN COMPLETE")
"""


def test_packer_is_clean_under_address_sanitizer(tmp_path):
    """libnpde_io.so's host code (driver/io_capi.cpp, driver/h5io.cpp) built with -fsanitize=address by g++
    and driven through npde.dataset's own calls: no invalid access, use after free or overflow."""
    gxx = shutil.which("g++")
    hdf5_inc = Path(os.environ.get("HDF5", "/opt/conda")) / "include"
    hdf5_lib = IO_LIB.parent / "hdf5" / "libhdf5.so.103"
    if not gxx or not (hdf5_inc / "hdf5.h").exists() or not hdf5_lib.exists():
        pytest.skip("needs g++, the HDF5 headers and the library build's HDF5 links")
    asan = subprocess.run([gxx, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("no AddressSanitizer runtime")
    drv = ROOT / "neural-pde-surrogate_amd" / "driver"
    lib = tmp_path / "libnpde_io_asan.so"
    subprocess.run([gxx, "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-std=c++17",
                    "-ffp-contract=off", "-DNPDE_WITH_HDF5", f"-I{hdf5_inc}", "-shared", "-fPIC", "-o", str(lib),
                    str(drv / "io_capi.cpp"), str(drv / "h5io.cpp"), str(hdf5_lib),
                    f"-Wl,-rpath,{hdf5_lib.parent}"], check=True)
    env = dict(os.environ, LD_PRELOAD=preload(asan), NPDE_IO_LIB=str(lib),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([sys.executable, "-c", ASAN_SCRIPT, str(ROOT / "neural-pde-surrogate_amd")], env=env,
                       capture_output=True, text=True, timeout=300)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and "PACKER RUN COMPLETE" in r.stdout, r.stderr[-4000:]


# The oracle's entry points on every built-in problem, scheme and limiter, on grids narrower than the stencil,
# on point-wise arrays, and through its non-physical failure path.
ORACLE_SCRIPT = r"""
import sys
from pathlib import Path
import numpy as np
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[2])
import oracle2d
from npde.families import riemann
orc = oracle2d.Oracle(Path(sys.argv[3]))
for test, n in (("lax_liu3", 64), ("sod_x", 40), ("sod_y", 40), ("sine2d", 32), ("kh", 32)):
    g, t_end, U = orc.setup(test, n)
    for s in ("rusanov", "hllc"):
        for l in ("none", "unlimited", "minmod", "vanleer"):
            a = U.copy()
            orc.run(g, s, l, 0.5, a, min(t_end, 0.05), nout=3, sizes=((n // 2,) if g.nx == g.ny else ()),
                    record_dts=True)
for nx, ny in ((1, 1), (1, 5), (5, 1), (2, 3), (7, 3)):
    for bx in ("periodic", "transmissive"):
        g = oracle2d.Grid.make(nx, ny, y=(0.0, ny / nx), bc_x=bx, bc_y="periodic")
        U = np.ascontiguousarray(riemann(3, 8)[0][:, :ny, :nx])
        orc.run(g, "hllc", "vanleer", 0.5, U, 0.02, nout=2, record_dts=True)
        orc.rhs(g, "hllc", "vanleer", U)
        orc.compute_dt(g, 0.5, U)
W = np.random.default_rng(0).uniform(0.1, 2.0, (1000, 4))
C = orc.to_conserved(W)
orc.to_primitive(C); orc.sound_speed(W); orc.physical_flux(C)
for l in ("none", "unlimited", "minmod", "vanleer"):
    orc.limiter(l, W[:, 0] - 1.0, W[:, 1] - 1.0)
for s in ("rusanov", "hllc"):
    for d in (0, 1):
        orc.face_flux(s, d, C, C[::-1].copy())
g, t_end, U = orc.setup("sod_x", 100)
assert orc.run(g, "rusanov", "unlimited", 0.5, U, 0.2, nout=1)["rc"] == 1
print("ORACLE RUN COMPLETE")
"""


def test_oracle_is_clean_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    """The checker itself (oracle/euler2d_oracle.cpp) built by g++ with -fsanitize=address,undefined and driven
    through oracle2d in a child process: no invalid access and no undefined behaviour."""
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("needs g++")
    rt = [subprocess.run([gxx, f"-print-file-name={lib}"], capture_output=True, text=True).stdout.strip()
          for lib in ("libasan.so", "libubsan.so")]
    if not all(os.path.isabs(p) and os.path.exists(p) for p in rt):
        pytest.skip("no sanitizer runtimes")
    lib = tmp_path / "liboracle2d_san.so"
    subprocess.run([gxx, "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
                    "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-shared", "-o", str(lib),
                    str(ROOT / "oracle" / "euler2d_oracle.cpp")], check=True)
    env = dict(os.environ, LD_PRELOAD=preload(*rt), ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-c", ORACLE_SCRIPT, str(ROOT / "oracle"),
                        str(ROOT / "neural-pde-surrogate_amd"), str(lib)], env=env, capture_output=True, text=True,
                       timeout=600)
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and "ORACLE RUN COMPLETE" in r.stdout, r.stderr[-4000:]

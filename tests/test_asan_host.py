"""Host code of the dataset job under AddressSanitizer, on the CPU: the HDF5 packer (libnpde_io.so, built here
by g++ with the sanitizer) driven by npde.dataset's own write paths in a child process."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
IO_LIB = ROOT / "neural-pde-surrogate_amd" / "lib" / "libnpde_io.so"


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
    env = dict(os.environ, LD_PRELOAD=asan, NPDE_IO_LIB=str(lib),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([sys.executable, "-c", ASAN_SCRIPT, str(ROOT / "neural-pde-surrogate_amd")], env=env,
                       capture_output=True, text=True, timeout=300)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and "PACKER RUN COMPLETE" in r.stdout, r.stderr[-4000:]

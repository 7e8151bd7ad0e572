"""The driver's HDF5 snapshot and initial-condition code (neural-pde-surrogate_amd/driver/h5io.cpp) on the CPU,
without a GPU: a small harness built here by g++ (with AddressSanitizer where its runtime exists) writes the
snapshot of the reference's own I/O test and reads files back, and h5dump checks them.

- The reference's test_io2d.cpp writes a snapshot with known values on a non-square, non-unit domain and
  tests/check_snapshot.py checks every value, shape and attribute: the same snapshot and checks here.
- A snapshot with the attributes of the reference CLI's 32^2 sine2d snapshot, and a trajectory file of the
  oracle's Lax-Liu 32^2 frames (equal to the reference CLI's), have the storage and attribute layouts of the
  reference CLI's files (h5dump -p -A, tests/golden/h5_layout.json; compressed sizes included) character
  for character.
- Reading a snapshot back as an initial condition (io_hdf5.hpp:171-175) gives the grid and the conserved
  state that to_conserved (state2d.hpp:16-19, the oracle's restatement) makes of its primitive datasets,
  bit for bit; files that are not initial conditions are refused with a message."""
import json
import os
import shutil
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "neural-pde-surrogate_amd" / "driver"
HDF5_LINK = ROOT / "neural-pde-surrogate_amd" / "lib" / "hdf5" / "libhdf5.so.103"
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
GOLDEN = Path(__file__).resolve().parent / "golden"

HARNESS = r"""
#include <cstdio>
#include <exception>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "h5io.hpp"

namespace {
// to_conserved in the reference's operation order (state2d.hpp:16-19)
void put(std::vector<double>& U, size_t n, size_t k, double rho, double u, double v, double p) {
    U[k] = rho;
    U[n + k] = rho * u;
    U[2 * n + k] = rho * v;
    U[3 * n + k] = p / (1.4 - 1.0) + 0.5 * rho * (u * u + v * v);
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1], path = argv[2];
    try {
        if (mode == "io_test") {  // test_io2d.cpp: 6 x 4 cells on [0, 3] x [-1, 1]
            const npde_grid_desc g{6, 4, 0.0, 3.0, -1.0, 1.0, NPDE_BC_PERIODIC, NPDE_BC_TRANSMISSIVE};
            const size_t n = 24;
            std::vector<double> U(4 * n);
            for (int j = 0; j < g.ny; ++j)
                for (int i = 0; i < g.nx; ++i) put(U, n, size_t(j) * g.nx + i, 1.0 + i, 0.1 * i, 0.01 * j, 2.0 + j);
            npde_io::write_snapshot(path, g, U, 0.125, {"io_test", "hllc", "vanleer", 0.5, 42});
        } else if (mode == "sine32") {  // the attributes of the reference CLI's sine2d 32^2 snapshot 1 of 2
            const npde_grid_desc g{32, 32, 0.0, 1.0, 0.0, 1.0, NPDE_BC_PERIODIC, NPDE_BC_PERIODIC};
            const size_t n = 1024;
            std::vector<double> U(4 * n);
            for (size_t k = 0; k < n; ++k) put(U, n, k, 1.0 + 0.001 * k, 1.0, 1.0, 1.0);
            npde_io::write_snapshot(path, g, U, 0.5, {"sine2d", "hllc", "vanleer", 0.5, 149});
        } else if (mode == "traj" && argc == 6) {  // lax_liu3 at 32^2 to t = 0.3, levels 32 and 16
            // argv[3]: frame times then the levels' float32 frames, as raw bytes; argv[4]: frames; argv[5]: steps
            const npde_grid_desc g{32, 32, 0.0, 1.0, 0.0, 1.0, NPDE_BC_TRANSMISSIVE, NPDE_BC_TRANSMISSIVE};
            const int frames = std::stoi(argv[4]);
            std::ifstream in(argv[3], std::ios::binary);
            std::vector<double> times(frames);
            in.read(reinterpret_cast<char*>(times.data()), std::streamsize(frames * sizeof(double)));
            std::vector<npde_io::Level> levels;
            for (int m : {32, 16}) {
                npde_io::Level L{m, m, std::vector<float>(size_t(frames) * 4 * m * m)};
                in.read(reinterpret_cast<char*>(L.data.data()), std::streamsize(L.data.size() * sizeof(float)));
                levels.push_back(std::move(L));
            }
            if (!in) throw std::runtime_error("short frame file");
            npde_io::write_trajectory(path, g, times, levels, {"lax_liu3", "hllc", "vanleer", 0.5, std::stoi(argv[5])},
                                      0.3, 0.0);
        } else if (mode == "read" && argc == 4) {
            npde_grid_desc g{};
            std::vector<double> U;
            const npde_io::InitInfo info = npde_io::read_initial_condition(path, g, U);
            std::printf("%d %d %.17g %.17g %.17g %.17g %d %d %.17g %s\n", g.nx, g.ny, g.x_min, g.x_max, g.y_min,
                        g.y_max, g.bc_x, g.bc_y, info.t_end, info.test.c_str());
            std::ofstream(argv[3], std::ios::binary).write(reinterpret_cast<const char*>(U.data()),
                                                           std::streamsize(U.size() * sizeof(double)));
        } else {
            return 2;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    inc = Path(os.environ.get("HDF5", "/opt/conda")) / "include"
    if not gxx or not (inc / "hdf5.h").exists() or not HDF5_LINK.exists() or not H5DUMP:
        pytest.skip("needs g++, h5dump, the HDF5 headers and the library build's HDF5 links")
    d = tmp_path_factory.mktemp("h5io")
    (d / "harness.cpp").write_text(HARNESS)
    asan = subprocess.run([gxx, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    san = ["-fsanitize=address", "-fno-omit-frame-pointer"] if os.path.isabs(asan) and os.path.exists(asan) else []
    exe = d / "h5io_harness"
    subprocess.run([gxx, "-O1", "-g", *san, "-std=c++17", "-ffp-contract=off", "-DNPDE_WITH_HDF5", f"-I{inc}",
                    f"-I{DRIVER}", "-o", str(exe), str(d / "harness.cpp"), str(DRIVER / "h5io.cpp"), str(HDF5_LINK),
                    f"-Wl,-rpath,{HDF5_LINK.parent}"], check=True)

    def run(*args):
        r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=120,
                           env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1"))
        assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
        return r
    return run


def h5(path, name, dtype):
    with tempfile.NamedTemporaryFile(suffix=".bin") as tmp:
        subprocess.run([H5DUMP, "-b", "LE", "-d", name, "-o", tmp.name, str(path)], check=True,
                       stdout=subprocess.DEVNULL)
        return np.fromfile(tmp.name, dtype=dtype)


def attr_text(path, name):
    out = subprocess.run([H5DUMP, "-a", name, str(path)], capture_output=True, text=True, check=True).stdout
    return out.split("DATA {", 1)[1].split("}", 1)[0].strip()


def test_reference_io_test_snapshot(harness, tmp_path):
    """tests/check_snapshot.py:17-43 on the snapshot test_io2d.cpp writes."""
    path = tmp_path / "io_test.h5"
    assert harness("io_test", path).returncode == 0
    nx, ny = 6, 4
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    rho, u, v, p = (h5(path, q, "<f8") for q in ("rho", "u", "v", "p"))
    assert all(a.size == nx * ny for a in (rho, u, v, p))
    assert "( 4, 6 )" in subprocess.run([H5DUMP, "-H", "-d", "rho", str(path)], capture_output=True, text=True).stdout
    rho, u, v, p = (a.reshape(ny, nx) for a in (rho, u, v, p))
    assert np.allclose(rho, 1.0 + i) and np.allclose(u, 0.1 * i) and np.allclose(v, 0.01 * j)
    assert np.allclose(p, 2.0 + j) and (rho > 0).all() and (p > 0).all()
    assert np.allclose(h5(path, "x", "<f8"), 0.25 + 0.5 * np.arange(nx))
    assert np.allclose(h5(path, "y", "<f8"), -0.75 + 0.5 * np.arange(ny))
    want = {"t": "(0): 0.125", "step": "(0): 42", "gamma": "(0): 1.4", "cfl": "(0): 0.5", "nx": "(0): 6",
            "ny": "(0): 4", "x_min": "(0): 0", "x_max": "(0): 3", "y_min": "(0): -1", "y_max": "(0): 1",
            "bc_x": '(0): "periodic"', "bc_y": '(0): "transmissive"', "test": '(0): "io_test"',
            "scheme": '(0): "hllc"', "limiter": '(0): "vanleer"'}
    for name, text in want.items():
        assert attr_text(path, name) == text, name


def test_snapshot_layout_matches_reference_cli(harness, tmp_path):
    """h5dump -p -A of a 32^2 snapshot with the attributes of the reference CLI's sine2d snapshot (datasets,
    types, shapes, contiguous storage and offsets, filters, fill and allocation properties, attribute types
    and values) equals the reference's, as tests/golden/make_golden.py recorded it."""
    gold = json.loads((GOLDEN / "h5_layout.json").read_text())["snap_sine32"]
    path = tmp_path / "snap_0001.h5"
    assert harness("sine32", path).returncode == 0
    got = subprocess.run([H5DUMP, "-p", "-A", str(path)], capture_output=True, text=True, check=True).stdout
    assert got.replace(str(path), path.name) == gold


def test_trajectory_layout_matches_reference_cli(harness, oracle, tmp_path):
    """The oracle's Lax-Liu 32^2 trajectory (frames at 32^2 and 16^2, which equal the reference CLI's bit for
    bit) written by write_trajectory: h5dump -p -A (datasets, float32 chunks of one frame, shuffle + deflate 4,
    the compressed sizes, attribute types and values) equals the reference CLI's --traj file
    (trajectory.hpp:53-91), and the frames read back unchanged."""
    gold = json.loads((GOLDEN / "h5_layout.json").read_text())["traj_ll32"]
    g, t_end, U = oracle.setup("lax_liu3", 32)
    out = oracle.run(g, "hllc", "vanleer", 0.5, U, t_end, nout=2, sizes=(32, 16))
    raw = tmp_path / "frames.bin"
    raw.write_bytes(out["frame_t"].astype("<f8").tobytes() + out["frames"][32].astype("<f4").tobytes()
                    + out["frames"][16].astype("<f4").tobytes())
    path = tmp_path / "ll32.h5"
    r = harness("traj", path, raw, 3, out["steps"])
    assert r.returncode == 0, r.stderr
    got = subprocess.run([H5DUMP, "-p", "-A", str(path)], capture_output=True, text=True, check=True).stdout
    assert got.replace(str(path), path.name) == gold
    for m in (32, 16):
        assert np.array_equal(h5(path, f"U_{m}", "<f4").reshape(3, 4, m, m).view(np.uint32),
                              out["frames"][m].view(np.uint32)), m
    assert np.array_equal(h5(path, "t", "<f8"), out["frame_t"])


def test_snapshot_reads_back_as_an_initial_condition(harness, oracle, tmp_path):
    """io_hdf5.hpp:171-175: a snapshot is a valid initial condition.  The grid comes back as written (no t_end
    in a snapshot: the driver then asks for --tend), and the state is to_conserved of the file's
    primitives, bit for bit."""
    path, raw = tmp_path / "io_test.h5", tmp_path / "U.bin"
    assert harness("io_test", path).returncode == 0
    r = harness("read", path, raw)
    assert r.returncode == 0, r.stderr
    f = r.stdout.split()
    assert [int(f[0]), int(f[1]), *map(float, f[2:6]), int(f[6]), int(f[7])] == [6, 4, 0.0, 3.0, -1.0, 1.0, 1, 0]
    assert float(f[8]) < 0.0 and f[9] == "io_test"
    U = np.fromfile(raw, dtype="<f8").reshape(4, 4, 6)
    W = np.stack([h5(path, q, "<f8").reshape(4, 6) for q in ("rho", "u", "v", "p")], axis=-1)
    want = oracle.to_conserved(W)  # (4, 6, 4): conserved per cell
    assert np.array_equal(U.view(np.uint64), np.moveaxis(want, -1, 0).view(np.uint64))


def test_files_that_are_not_initial_conditions_are_refused(harness, tmp_path):
    """A missing path, an empty file, a non-HDF5 file and an HDF5 file without the snapshot attributes each
    end in a message (the driver prints it and exits with code 1)."""
    (tmp_path / "empty.h5").write_bytes(b"")
    (tmp_path / "junk.h5").write_bytes(bytes(range(256)) * 4)
    cases = {"missing.h5": "could not open", "empty.h5": "could not open", "junk.h5": "could not open"}
    for name, msg in cases.items():
        r = harness("read", tmp_path / name, tmp_path / "U.bin")
        assert r.returncode == 1 and msg in r.stderr, (name, r.stderr)
    # an HDF5 file without the snapshot attributes: a packed dataset file
    io_lib = ROOT / "neural-pde-surrogate_amd" / "lib" / "libnpde_io.so"
    if io_lib.exists():
        from npde.dataset import write_packed
        write_packed(tmp_path / "packed.h5", np.ones((1, 2, 4, 4, 4), np.float32), np.zeros((1, 2)), ["riemann"],
                     [0], "train", 0.1)
        r = harness("read", tmp_path / "packed.h5", tmp_path / "U.bin")
        assert r.returncode == 1 and "missing attribute nx" in r.stderr, r.stderr

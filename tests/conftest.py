import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "neural-pde-surrogate_amd"
GOLDEN = Path(__file__).resolve().parent / "golden"
for p in (ROOT, PKG, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    import oracle2d
    if not oracle2d.ORACLE_SO.exists():
        oracle2d.build()
    return oracle2d.Oracle()


@pytest.fixture(scope="session")
def reference():
    import oracle2d
    if not oracle2d.have_reference():
        pytest.skip("reference build (oracle/_ref) not present on this machine")
    return oracle2d.Reference()


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(GOLDEN / name, allow_pickle=False)
    return load


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def npde_mod():
    if not gpu_available():
        pytest.skip("no GPU")
    import npde
    return npde


def bitwise_equal(a, b) -> bool:
    """Same shape, dtype and bit pattern element by element (so +0 != -0); NaNs match NaNs at the same
    positions whatever their payload (failed states hold NaNs)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind != "f":
        return np.array_equal(a, b)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    u = {8: np.uint64, 4: np.uint32, 2: np.uint16}[a.itemsize]
    return np.array_equal(a.view(u)[~na], b.view(u)[~nb])


def run_case_grid(arr):
    """Grid from a fixture's (nx, ny, x_min, x_max, y_min, y_max, bc_x, bc_y) vector."""
    import oracle2d
    nx, ny, x0, x1, y0, y1, bx, by = arr
    return oracle2d.Grid(int(nx), int(ny), x0, x1, y0, y1, int(bx), int(by))


os.environ.setdefault("OMP_NUM_THREADS", str(min(8, os.cpu_count() or 1)))

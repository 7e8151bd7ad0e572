"""ctypes binding of libnpde.so (the C-ABI in include/npde.h).

The library is the only compute path: importing this module when the HIP library has not been
built raises immediately, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent          # neural-pde-surrogate_amd/
LIB_PATH = Path(os.environ.get("NPDE_LIB", PKG_ROOT / "lib" / "libnpde.so"))

OK, ERR_INVALID, ERR_NONPHYSICAL, ERR_DEVICE, ERR_NOMEM, ERR_STATE = range(6)
LAYOUT_PLANES, LAYOUT_GHOSTED = 0, 1


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("x_min", C.c_double), ("x_max", C.c_double),
                ("y_min", C.c_double), ("y_max", C.c_double), ("bc_x", C.c_int), ("bc_y", C.c_int)]


class BadCell(C.Structure):
    _fields_ = [("instance", C.c_int), ("i", C.c_int), ("j", C.c_int), ("rho", C.c_double), ("p", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("rows_per_task", C.c_int), ("waves_per_block", C.c_int), ("steps_per_graph", C.c_int),
                ("use_graph", C.c_int), ("row_groups", C.c_int), ("seams", C.c_int), ("waves_per_simd", C.c_int),
                ("net_period", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
_H = C.c_void_p

# (name, restype, argtypes)
_SIGNATURES = [
    ("npde_create", C.c_int, [C.POINTER(GridDesc), C.c_int, C.c_int, C.c_double, C.POINTER(_H)]),
    ("npde_batch_create", C.c_int, [C.c_int, C.POINTER(GridDesc), C.c_int, C.c_int, C.c_double, C.POINTER(_H)]),
    ("npde_destroy", C.c_int, [_H]),
    ("npde_num_instances", C.c_int, [_H]),
    ("npde_last_error", C.c_char_p, [_H]),
    ("npde_set_state", C.c_int, [_H, C.c_int, _D, C.c_int]),
    ("npde_get_state", C.c_int, [_H, C.c_int, _D, C.c_int]),
    ("npde_set_state_device", C.c_int, [_H, C.c_int, C.c_void_p]),
    ("npde_get_state_device", C.c_int, [_H, C.c_int, C.c_void_p]),
    ("npde_get_time", C.c_int, [_H, C.c_int, _D, C.POINTER(C.c_longlong)]),
    ("npde_get_times", C.c_int, [_H, _D, C.POINTER(C.c_longlong)]),
    ("npde_state_pointer", C.c_int, [_H, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_longlong),
                                     C.POINTER(C.c_longlong)]),
    ("npde_compute_dt", C.c_int, [_H, _D, C.POINTER(BadCell)]),
    ("npde_step", C.c_int, [_H, _D]),
    ("npde_advance_to", C.c_int, [_H, C.c_double, C.POINTER(BadCell)]),
    ("npde_advance_steps", C.c_int, [_H, C.c_longlong, C.c_double, C.POINTER(BadCell)]),
    ("npde_time_stages", C.c_int, [_H, C.c_int, _D, _D]),
    ("npde_run_trajectory", C.c_int, [_H, C.c_double, C.c_double, C.c_int, _I, C.c_int, C.POINTER(BadCell)]),
    ("npde_get_trajectory", C.c_int, [_H, C.c_int, C.c_int, _F]),
    ("npde_get_frame_times", C.c_int, [_H, C.c_int, _D]),
    ("npde_block_average", C.c_int, [_H, C.c_int, C.c_int, _F]),
    ("npde_block_average_noalloc", C.c_int, [_H, C.c_int, C.c_int, _F]),
    ("npde_set_flux_correction", C.c_int, [_H, C.c_int, C.c_void_p, C.c_void_p]),
    ("npde_get_tuning", C.c_int, [_H, C.POINTER(Tuning)]),
    ("npde_set_tuning", C.c_int, [_H, C.POINTER(Tuning)]),
    ("npde_autotune", C.c_int, [_H, C.c_char_p, C.POINTER(Tuning)]),
    ("npde_set_arithmetic", C.c_int, [_H, C.c_int]),
    ("npde_get_arithmetic", C.c_int, [_H, C.POINTER(C.c_int)]),
    ("npde_get_task_layout", C.c_int, [_H, C.POINTER(C.c_int)]),
    ("npde_get_occupancy", C.c_int, [_H, C.POINTER(C.c_int)]),
    ("npde_set_profiling", C.c_int, [_H, C.c_int]),
    ("npde_get_profile", C.c_int, [_H, _D, C.POINTER(C.c_longlong)]),
    ("npde_last_elapsed_ms", C.c_int, [_H, _D]),
    ("npde_stream", C.c_void_p, [_H]),
    ("npde_builtin_problem", C.c_int, [C.c_char_p, C.c_int, C.POINTER(GridDesc), _D, _D]),
    ("npde_selftest_arith", C.c_int, [_D, _D, C.c_longlong, _D, _D, _D]),
    ("npde_conv3x3", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                               C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    ("npde_conv3x3_wgrad", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("npde_mfma_probe", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("npde_mfma32_probe", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("npde_set_flux_network", C.c_int, [_H, C.c_void_p, C.c_void_p, C.c_double]),
    ("npde_set_flux_network_limit", C.c_int, [_H, C.c_double]),
    ("npde_flux_network_switched_off", C.c_int, [_H, _I]),
    ("npde_flux_network_eval", C.c_int, [_H]),
    ("npde_flux_network_peak", C.c_int, [_H, _F, C.c_int]),
    ("npde_flux_network_planes", C.c_int, [_H, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
]

EXPORTED = [name for name, _, _ in _SIGNATURES]

_lib = None


def load(path: Path | str | None = None) -> C.CDLL:
    """Load libnpde.so once; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"libnpde.so not found at {p}: build it first "
                           f"(python -c 'import __graft_entry__ as g; g.build()' or make -C {PKG_ROOT / 'csrc'})")
    lib = C.CDLL(str(p))
    for name, res, args in _SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class NpdeError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class NonPhysicalStateError(NpdeError):
    """rho <= 0 or p <= 0 where a step was needed (the reference exits(1) there, solver2d.hpp:116-126)."""

    def __init__(self, status, message, bad: BadCell):
        super().__init__(status, message)
        self.instance, self.i, self.j, self.rho, self.p = bad.instance, bad.i, bad.j, bad.rho, bad.p


def check(handle, status: int, bad: BadCell | None = None):
    if status == OK:
        return
    msg = load().npde_last_error(handle)
    msg = msg.decode() if msg else f"npde status {status}"
    if status == ERR_NONPHYSICAL and bad is not None:
        raise NonPhysicalStateError(status, msg, bad)
    raise NpdeError(status, msg)

"""Python mirror of the reference's solver interface, running on the HIP library.

Reference interface (solver/src):
  grid2d.hpp:9-39     BC, Grid (nx, ny, ng=2, bounds, dx, dy, idx, xc, yc), make_grid
  grid2d.hpp:43-66    apply_bcs
  riemann2d.hpp:8     FluxFn  (rusanov_flux | hllc_flux), chosen by name in main2d.cpp:145-148
  limiters.hpp:5      Limiter (none | unlimited | minmod | vanleer), main2d.cpp:150-155
  solver2d.hpp:57-103 Solver(grid, flux, lim, cfl).compute_dt(U), .step(U, dt)

State arrays are numpy float64 in either layout:
  planes   (4, ny, nx)          rho, rho*u, rho*v, E over the interior cells
  ghosted  (ny+4, nx+4, 4)      the reference's std::vector<Conserved> (ng = 2), row-major in j
A non-physical state raises NonPhysicalStateError carrying the first bad cell (i, j) in row-major
order and its rho and p -- the reference prints the same and exits(1) (solver2d.hpp:116-126).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import BadCell, GridDesc, Tuning, check

TRANSMISSIVE, PERIODIC = 0, 1
BC_NAMES = {"transmissive": TRANSMISSIVE, "periodic": PERIODIC}
FLUXES = {"rusanov": 0, "hllc": 1}
LIMITERS = {"none": 0, "unlimited": 1, "minmod": 2, "vanleer": 3}
GAMMA = 1.4


def _bc(v):
    return BC_NAMES[v] if isinstance(v, str) else int(v)


@dataclass
class Grid:
    """Uniform grid with ng = 2 ghost layers (grid2d.hpp:12-28)."""
    nx: int
    ny: int
    x_min: float = 0.0
    x_max: float = 1.0
    y_min: float = 0.0
    y_max: float = 1.0
    bc_x: int = TRANSMISSIVE
    bc_y: int = TRANSMISSIVE
    ng: int = 2

    @property
    def dx(self) -> float:
        return (self.x_max - self.x_min) / self.nx

    @property
    def dy(self) -> float:
        return (self.y_max - self.y_min) / self.ny

    def nx_tot(self) -> int:
        return self.nx + 2 * self.ng

    def ny_tot(self) -> int:
        return self.ny + 2 * self.ng

    def size(self) -> int:
        return self.nx_tot() * self.ny_tot()

    def idx(self, i: int, j: int) -> int:
        return (j + self.ng) * self.nx_tot() + (i + self.ng)

    def xc(self, i):
        return self.x_min + (np.asarray(i) + 0.5) * self.dx

    def yc(self, j):
        return self.y_min + (np.asarray(j) + 0.5) * self.dy

    def desc(self) -> GridDesc:
        return GridDesc(self.nx, self.ny, self.x_min, self.x_max, self.y_min, self.y_max, self.bc_x, self.bc_y)

    @classmethod
    def from_desc(cls, d: GridDesc) -> "Grid":
        return cls(d.nx, d.ny, d.x_min, d.x_max, d.y_min, d.y_max, d.bc_x, d.bc_y)


def make_grid(nx, ny, x_min=0.0, x_max=1.0, y_min=0.0, y_max=1.0, bc_x=TRANSMISSIVE, bc_y=TRANSMISSIVE) -> Grid:
    """grid2d.hpp:30-39."""
    if nx <= 0 or ny <= 0:
        raise ValueError("nx and ny must be positive")
    return Grid(int(nx), int(ny), float(x_min), float(x_max), float(y_min), float(y_max), _bc(bc_x), _bc(bc_y))


def apply_bcs(grid: Grid, U: np.ndarray) -> np.ndarray:
    """Fill the ghost ring of a ghosted (ny+4, nx+4, 4) array in place: x ghosts on the physical rows,
    then y ghosts over all columns, so corners are filled (grid2d.hpp:43-66)."""
    g, ng = grid, grid.ng
    A = U.reshape(g.ny_tot(), g.nx_tot(), 4)
    rows = slice(ng, ng + g.ny)
    for k in range(1, ng + 1):
        src_l = g.nx - k if g.bc_x == PERIODIC else 0
        src_r = k - 1 if g.bc_x == PERIODIC else g.nx - 1
        A[rows, ng - k] = A[rows, ng + src_l]
        A[rows, ng + g.nx - 1 + k] = A[rows, ng + src_r]
    for k in range(1, ng + 1):
        src_b = g.ny - k if g.bc_y == PERIODIC else 0
        src_t = k - 1 if g.bc_y == PERIODIC else g.ny - 1
        A[ng - k, :] = A[ng + src_b, :]
        A[ng + g.ny - 1 + k, :] = A[ng + src_t, :]
    return U


def _layout(grid: Grid, U: np.ndarray) -> int:
    if U.dtype != np.float64 or not U.flags.c_contiguous:
        raise TypeError("state must be a C-contiguous float64 array")
    if U.shape == (4, grid.ny, grid.nx):
        return _lib.LAYOUT_PLANES
    if U.size == grid.size() * 4 and (U.ndim == 1 or U.shape[-1] == 4):
        return _lib.LAYOUT_GHOSTED
    raise ValueError(f"state of shape {U.shape} matches neither (4, {grid.ny}, {grid.nx}) "
                     f"nor ({grid.ny_tot()}, {grid.nx_tot()}, 4)")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Batch:
    """Many independent instances (any sizes) stepped together on the GPU, each with its own dt, time
    and frame landing.  `Solver` is the one-instance case with the reference's method names."""

    def __init__(self, grids, flux="hllc", limiter="vanleer", cfl=0.5):
        lib = _lib.load()
        self.grids = [g if isinstance(g, Grid) else Grid.from_desc(g) for g in grids]
        if not self.grids:
            raise ValueError("empty batch")
        self.flux = flux
        self.limiter = limiter
        self.cfl = float(cfl)
        if flux not in FLUXES:
            raise ValueError(f"Unknown scheme '{flux}' (use rusanov or hllc)")
        if limiter not in LIMITERS:
            raise ValueError(f"Unknown limiter '{limiter}'")
        if not self.cfl > 0.0:
            raise ValueError("cfl must be positive")
        arr = (GridDesc * len(self.grids))(*[g.desc() for g in self.grids])
        h = C.c_void_p()
        st = lib.npde_batch_create(len(self.grids), arr, FLUXES[flux], LIMITERS[limiter], self.cfl, C.byref(h))
        if st != _lib.OK:
            raise _lib.NpdeError(st, f"npde_batch_create failed with status {st}")
        self._h = h
        self._lib = lib

    # -- lifetime --
    def close(self):
        if getattr(self, "_h", None):
            self._lib.npde_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n(self) -> int:
        return len(self.grids)

    # -- state --
    def set_state(self, inst: int, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.float64)
        check(self._h, self._lib.npde_set_state(self._h, inst, _dptr(U), _layout(self.grids[inst], U)))

    def get_state(self, inst: int, ghosted: bool = False) -> np.ndarray:
        g = self.grids[inst]
        if ghosted:
            out = np.empty((g.ny_tot(), g.nx_tot(), 4))
            check(self._h, self._lib.npde_get_state(self._h, inst, _dptr(out), _lib.LAYOUT_GHOSTED))
        else:
            out = np.empty((4, g.ny, g.nx))
            check(self._h, self._lib.npde_get_state(self._h, inst, _dptr(out), _lib.LAYOUT_PLANES))
        return out

    def set_state_device(self, inst: int, ptr: int):
        """ptr: device address of a contiguous (4, ny, nx) float64 buffer (e.g. tensor.data_ptr())."""
        check(self._h, self._lib.npde_set_state_device(self._h, inst, C.c_void_p(ptr)))

    def get_state_device(self, inst: int, ptr: int):
        check(self._h, self._lib.npde_get_state_device(self._h, inst, C.c_void_p(ptr)))

    def state_view(self):
        """Zero-copy torch view (n, 4, ny, nx) of every instance's state in device memory (all instances
        must have the same size).  Read it after a call returns; write it only between calls, and order
        the write before the next call: the library's stream does not wait for torch's streams, so issue
        the write under torch.cuda.stream(torch.cuda.ExternalStream(batch.stream)), or call
        torch.cuda.synchronize() after it.  The view aliases the library's buffer: it is valid until the
        batch is closed."""
        import torch
        g0 = self.grids[0]
        if any((g.nx, g.ny) != (g0.nx, g0.ny) for g in self.grids):
            raise ValueError("state_view needs instances of one size")
        ptr, plane, stride = C.c_void_p(), C.c_longlong(), C.c_longlong()
        check(self._h, self._lib.npde_state_pointer(self._h, 0, C.byref(ptr), C.byref(plane), C.byref(stride)))
        if self.n > 1:
            p1, s1 = C.c_void_p(), C.c_longlong()
            check(self._h, self._lib.npde_state_pointer(self._h, 1, C.byref(p1), C.byref(s1), None))
            stride = C.c_longlong((p1.value - ptr.value) // 8)

        class _View:
            __cuda_array_interface__ = {
                "shape": (self.n, 4, g0.ny, g0.nx), "typestr": "<f8", "data": (ptr.value, False), "version": 3,
                "strides": (stride.value * 8, plane.value * 8, g0.nx * 8, 8)}

        view = torch.as_tensor(_View(), device="cuda")
        view._npde_owner = self  # keeps the batch (and its device buffer) alive while the view is
        return view

    def time(self, inst: int = 0):
        t, steps = C.c_double(), C.c_longlong()
        check(self._h, self._lib.npde_get_time(self._h, inst, C.byref(t), C.byref(steps)))
        return t.value, steps.value

    def times(self):
        """(t, steps) arrays of every instance, read in one device-to-host copy."""
        t = np.empty(self.n)
        steps = np.empty(self.n, dtype=np.int64)
        check(self._h, self._lib.npde_get_times(self._h, t.ctypes.data_as(C.POINTER(C.c_double)),
                                                 steps.ctypes.data_as(C.POINTER(C.c_longlong))))
        return t, steps

    # -- hot path --
    def compute_dts(self) -> np.ndarray:
        dts = np.empty(self.n)
        bad = BadCell()
        check(self._h, self._lib.npde_compute_dt(self._h, _dptr(dts), C.byref(bad)), bad)
        return dts

    def step_all(self, dts):
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, dtype=np.float64), (self.n,)))
        check(self._h, self._lib.npde_step(self._h, _dptr(dts)))

    def advance_to(self, t_target: float):
        bad = BadCell()
        check(self._h, self._lib.npde_advance_to(self._h, float(t_target), C.byref(bad)), bad)

    def advance_steps(self, nsteps: int, t_target: float):
        """Device loop toward t_target that stops every instance after nsteps more steps."""
        bad = BadCell()
        check(self._h, self._lib.npde_advance_steps(self._h, int(nsteps), float(t_target), C.byref(bad)), bad)

    def time_stages(self, nsteps: int):
        """Average device ms of the stage-1 and stage-2 kernels over the next nsteps CFL steps, launched
        one at a time with HIP events around each (the state advances)."""
        a, b = C.c_double(), C.c_double()
        check(self._h, self._lib.npde_time_stages(self._h, int(nsteps), C.byref(a), C.byref(b)))
        return a.value, b.value

    def run_trajectory(self, t_end: float, nout: int, sizes=(), t_start: float = 0.0):
        """Pre-roll to t_start, then nout frames landing on t_start + (t_end - t_start) k / nout
        (main2d.cpp:221-246), recording block-averaged float32 frames at `sizes`."""
        sizes = [int(m) for m in sizes]
        arr = (C.c_int * max(1, len(sizes)))(*sizes)
        bad = BadCell()
        # set first: when an instance fails, the others' frames are complete and can still be read
        self._traj = (sizes, int(nout))
        check(self._h, self._lib.npde_run_trajectory(self._h, float(t_start), float(t_end), int(nout), arr,
                                                     len(sizes), C.byref(bad)), bad)

    def trajectory(self, inst: int, m: int) -> np.ndarray:
        sizes, nout = self._traj
        g = self.grids[inst]
        f = g.nx // m
        out = np.empty((nout + 1, 4, g.ny // f, m), dtype=np.float32)
        check(self._h, self._lib.npde_get_trajectory(self._h, inst, m, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def frame_times(self, inst: int) -> np.ndarray:
        _, nout = self._traj
        out = np.empty(nout + 1)
        check(self._h, self._lib.npde_get_frame_times(self._h, inst, _dptr(out)))
        return out

    def block_average(self, inst: int, m: int) -> np.ndarray:
        g = self.grids[inst]
        f = g.nx // m if m > 0 else 0
        if m <= 0 or g.nx % m or g.ny % f:
            raise ValueError(f"save size {m} does not divide the grid")
        out = np.empty((4, g.ny // f, m), dtype=np.float32)
        check(self._h, self._lib.npde_block_average(self._h, inst, m, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_flux_correction(self, inst: int, corr_x_ptr: int | None, corr_y_ptr: int | None):
        check(self._h, self._lib.npde_set_flux_correction(
            self._h, inst, C.c_void_p(corr_x_ptr) if corr_x_ptr else None,
            C.c_void_p(corr_y_ptr) if corr_y_ptr else None))

    def set_flux_network(self, model=None):
        """Run the learned flux-correction network (a FluxCorrector of the 8 -> 4 x 32 -> 8 stack) inside
        the device loop: npde_run_trajectory evaluates it at every frame landing of every instance, from the
        state at that time, and the stage kernels add its scaled factors to the face fluxes until the next
        landing.  None switches it off.  The packed weights stay referenced by the batch."""
        if model is None:
            check(self._h, self._lib.npde_set_flux_network(self._h, None, None, 0.0))
            self._net = None
            return
        import torch
        frags, biases, scale = model.pack_fused()
        if frags.device.type != "cuda" or biases.device != frags.device:
            raise ValueError("set_flux_network needs the model on the GPU (the kernels read its packed weights "
                             "from device memory)")
        # packed by torch on its own stream, which the library's stream does not wait for: complete them
        # before the library can read them
        torch.cuda.current_stream(frags.device).synchronize()
        self._net = (frags, biases)  # device buffers the library reads until switched off
        check(self._h, self._lib.npde_set_flux_network(self._h, C.c_void_p(frags.data_ptr()),
                                                       C.c_void_p(biases.data_ptr()), scale))

    def set_flux_network_limit(self, ratio: float):
        """Switch the network off for an instance (zero factors: the plain coarse solver) at the first frame
        landing of a run where its CFL signal-speed sum exceeds `ratio` times its first frame's; 0 never
        (default 2.5)."""
        check(self._h, self._lib.npde_set_flux_network_limit(self._h, float(ratio)))

    def flux_network_switched_off(self) -> np.ndarray:
        """Per instance, True if the last run switched its network off."""
        out = np.zeros(self.n, dtype=np.int32)
        check(self._h, self._lib.npde_flux_network_switched_off(self._h, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out.astype(bool)

    def flux_network_eval(self):
        """Evaluate the network now on every instance's current state (its factors for the next steps)."""
        check(self._h, self._lib.npde_flux_network_eval(self._h))

    def flux_network_peak(self, reset: bool = False) -> np.ndarray:
        """Per instance, the largest density the network has seen at its evaluations since the last reset."""
        out = np.zeros(self.n, dtype=np.float32)
        check(self._h, self._lib.npde_flux_network_peak(self._h, out.ctypes.data_as(C.POINTER(C.c_float)),
                                                        int(bool(reset))))
        return out

    def flux_network_planes(self, inst: int):
        """(corr_x, corr_y) device addresses of the network's factor planes of an instance."""
        x, y = C.c_void_p(), C.c_void_p()
        check(self._h, self._lib.npde_flux_network_planes(self._h, inst, C.byref(x), C.byref(y)))
        return x.value, y.value

    # -- performance --
    @property
    def tuning(self) -> dict:
        t = Tuning()
        check(self._h, self._lib.npde_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning()
        check(self._h, self._lib.npde_get_tuning(self._h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(self._h, self._lib.npde_set_tuning(self._h, C.byref(t)))

    ARITH = {"exact": 0, "fast": 1}

    @property
    def arithmetic(self) -> str:
        a = C.c_int()
        check(self._h, self._lib.npde_get_arithmetic(self._h, C.byref(a)))
        return "fast" if a.value == 1 else "exact"

    def set_arithmetic(self, mode: str):
        """'exact' (default; bit-identical to the reference solver) or 'fast' (reciprocal divisions,
        shortened square root, fused multiply-adds: rounding-level differences per step)."""
        if mode not in self.ARITH:
            raise ValueError(f"arithmetic must be 'exact' or 'fast', not {mode!r}")
        check(self._h, self._lib.npde_set_arithmetic(self._h, self.ARITH[mode]))

    TASK_LAYOUTS = ("strips", "row_groups", "seams")

    @property
    def task_layout(self) -> str:
        """Task layout of the stage kernels under the current tuning: 'strips' (60 columns and a halo per
        wavefront), 'row_groups' or 'seams' (64 columns per wavefront, which solves its strip boundaries first)."""
        m = C.c_int()
        check(self._h, self._lib.npde_get_task_layout(self._h, C.byref(m)))
        return self.TASK_LAYOUTS[m.value]

    @property
    def occupancy(self) -> int:
        """Resident wavefronts per SIMD of the stage-2 kernel under the current tuning (registers, LDS with an
        occupancy cap's reservation, wavefront slots)."""
        m = C.c_int()
        check(self._h, self._lib.npde_get_occupancy(self._h, C.byref(m)))
        return m.value

    def autotune(self, cache_path: str | None = None) -> dict:
        t = Tuning()
        check(self._h, self._lib.npde_autotune(self._h, cache_path.encode() if cache_path else None, C.byref(t)))
        return t.as_dict()

    PROFILE_KINDS = ("stage1", "stage2", "finish", "frame", "wavespeed", "network")

    def set_profiling(self, on: bool = True):
        """Per-kernel profile: while on, kernels launch one at a time (no graphs) between HIP events (the profiled
        run is slower).  Switching on clears the accumulated times."""
        check(self._h, self._lib.npde_set_profiling(self._h, int(bool(on))))

    def profile(self) -> dict:
        """{kind: (device milliseconds, launches)} accumulated while profiling was on."""
        ms = (C.c_double * len(self.PROFILE_KINDS))()
        n = (C.c_longlong * len(self.PROFILE_KINDS))()
        check(self._h, self._lib.npde_get_profile(self._h, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(self.PROFILE_KINDS)}

    def last_elapsed_ms(self) -> float:
        ms = C.c_double()
        check(self._h, self._lib.npde_last_elapsed_ms(self._h, C.byref(ms)))
        return ms.value

    @property
    def stream(self) -> int:
        return int(self._lib.npde_stream(self._h) or 0)


class Solver(Batch):
    """e2d::Solver (solver2d.hpp:57-103) on the GPU.

    compute_dt(U) and step(U, dt) take the caller's array, as the reference does; step updates it in
    place.  For the ghosted layout the ghost ring ends as the reference leaves it: filled from the state
    at the start of the step (apply_bcs at the top of compute_rhs, solver2d.hpp:131).  The device-
    resident methods of Batch (set_state, advance_to, run_trajectory, ...) avoid the host copies."""

    def __init__(self, grid: Grid, flux="hllc", limiter="vanleer", cfl=0.5):
        super().__init__([grid], flux, limiter, cfl)
        self.grid = self.grids[0]

    def compute_dt(self, U: np.ndarray | None = None) -> float:
        if U is not None:
            self.set_state(0, U)
        return float(self.compute_dts()[0])

    def step(self, U: np.ndarray, dt: float) -> np.ndarray:
        layout = _layout(self.grid, U)
        self.set_state(0, U)
        self.step_all([dt])
        if layout == _lib.LAYOUT_PLANES:
            U[...] = self.get_state(0)
        else:
            apply_bcs(self.grid, U)  # ghosts of U^n, as the reference leaves them
            g, ng = self.grid, self.grid.ng
            A = U.reshape(g.ny_tot(), g.nx_tot(), 4)
            A[ng:ng + g.ny, ng:ng + g.nx, :] = np.moveaxis(self.get_state(0), 0, -1)
        return U


def builtin_problem(test: str, n: int):
    """Built-in initial condition of the driver (main2d.cpp:52-135): returns (grid, t_end, U planes)."""
    lib = _lib.load()
    d, t_end = GridDesc(), C.c_double()
    if lib.npde_builtin_problem(test.encode(), int(n), C.byref(d), C.byref(t_end), None) != _lib.OK:
        raise ValueError(f"Unknown test '{test}' (use sod_x, sod_y, lax_liu3, sine2d or kh)")
    U = np.zeros((4, d.ny, d.nx))
    lib.npde_builtin_problem(test.encode(), int(n), C.byref(d), C.byref(t_end), _dptr(U))
    return Grid.from_desc(d), t_end.value, U


def to_conserved(rho, u, v, p):
    """state2d.hpp:16-19 (same operation order), numpy arrays -> (4, ...) planes."""
    rho, u, v, p = (np.asarray(a, dtype=np.float64) for a in (rho, u, v, p))
    kin = (0.5 * rho) * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, p / (GAMMA - 1.0) + kin])


def to_primitive(U):
    """state2d.hpp:21-25, (4, ...) conserved planes -> (rho, u, v, p)."""
    rho, mx, my, E = U
    u = mx / rho
    v = my / rho
    p = (GAMMA - 1.0) * (E - (0.5 * rho) * (u * u + v * v))
    return rho, u, v, p

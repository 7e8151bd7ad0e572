"""ctypes access to the CPU oracle (liboracle2d.so) and, where built, the reference solver
(_ref/libref2d.so).  TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker -- never by the product package.

State arrays are float64 of shape (4, ny, nx): rho, rho*u, rho*v, E over the interior cells.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle2d.so"
REF_SO = HERE / "_ref" / "libref2d.so"
REF_EULER2D = HERE / "_ref" / "euler2d"

BC = {"transmissive": 0, "periodic": 1}
SCHEME = {"rusanov": 0, "hllc": 1}
LIMITER = {"none": 0, "unlimited": 1, "minmod": 2, "vanleer": 3}


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("x_min", C.c_double), ("x_max", C.c_double),
                ("y_min", C.c_double), ("y_max", C.c_double), ("bc_x", C.c_int), ("bc_y", C.c_int)]

    @classmethod
    def make(cls, nx, ny, x=(0.0, 1.0), y=(0.0, 1.0), bc_x="transmissive", bc_y="transmissive"):
        bx = BC[bc_x] if isinstance(bc_x, str) else int(bc_x)
        by = BC[bc_y] if isinstance(bc_y, str) else int(bc_y)
        return cls(nx, ny, x[0], x[1], y[0], y[1], bx, by)


_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
_L = C.POINTER(C.c_longlong)


def _p(a, t=_D):
    return None if a is None else a.ctypes.data_as(t)


def build():
    """Compile the oracle (and the reference harness when the reference sources exist)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE), "-j8"], check=True)


class _Lib:
    def __init__(self, path: Path, prefix: str):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (run: make -C {HERE})")
        self.lib = C.CDLL(str(path))
        self.prefix = prefix
        L = self.lib
        for name in ("to_primitive", "to_conserved", "physical_flux"):
            f = getattr(L, f"{prefix}_{name}")
            f.argtypes = [_D, _D, C.c_long]
            f.restype = None
        f = getattr(L, f"{prefix}_sound_speed")
        f.argtypes = [_D, _D, C.c_long]
        f = getattr(L, f"{prefix}_limiter")
        f.argtypes = [C.c_int, _D, _D, _D, C.c_long]
        f = getattr(L, f"{prefix}_face_flux")
        f.argtypes = [C.c_int, C.c_int, _D, _D, _D, C.c_long]

    def _rec(self, name, X, width_out=4):
        X = np.ascontiguousarray(X, dtype=np.float64)
        n = X.size // 4
        out = np.empty(n * width_out if width_out > 1 else n, dtype=np.float64)
        getattr(self.lib, f"{self.prefix}_{name}")(_p(X), _p(out), n)
        return out.reshape(X.shape) if width_out == 4 else out.reshape(X.shape[:-1])

    def to_primitive(self, U):
        return self._rec("to_primitive", U)

    def to_conserved(self, W):
        return self._rec("to_conserved", W)

    def physical_flux(self, U):
        return self._rec("physical_flux", U)

    def sound_speed(self, W):
        return self._rec("sound_speed", W, width_out=1)

    def limiter(self, kind, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.empty_like(a)
        getattr(self.lib, f"{self.prefix}_limiter")(LIMITER.get(kind, kind), _p(a), _p(b), _p(out), a.size)
        return out

    def face_flux(self, scheme, direction, UL, UR):
        UL = np.ascontiguousarray(UL, dtype=np.float64)
        UR = np.ascontiguousarray(UR, dtype=np.float64)
        F = np.empty_like(UL)
        getattr(self.lib, f"{self.prefix}_face_flux")(SCHEME.get(scheme, scheme), int(direction),
                                                      _p(UL), _p(UR), _p(F), UL.size // 4)
        return F


class Oracle(_Lib):
    """The repo's CPU restatement (C++17, OpenMP)."""

    def __init__(self, path: Path = ORACLE_SO):
        super().__init__(path, "orc")
        L = self.lib
        L.orc_rhs.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, _D, _D, C.c_int]
        L.orc_compute_dt.argtypes = [C.POINTER(Grid), C.c_double, _D, _D, _D, _D, _I, _I, C.c_int]
        L.orc_compute_dt.restype = C.c_int
        L.orc_step.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, _D, C.c_double, C.c_int]
        L.orc_block_average.argtypes = [C.POINTER(Grid), _D, C.c_int, _F]
        L.orc_run.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, C.c_double, _D, C.c_double, C.c_double,
                              C.c_int, _I, C.c_int, _F, _D, _D, C.c_longlong, _L, _D, _I, C.c_int]
        L.orc_run.restype = C.c_int
        L.orc_setup.argtypes = [C.c_char_p, C.c_int, C.POINTER(Grid), _D, _D]
        L.orc_setup.restype = C.c_int

    def setup(self, test: str, n: int):
        g = Grid()
        t_end = C.c_double()
        if self.lib.orc_setup(test.encode(), n, C.byref(g), C.byref(t_end), None):
            raise ValueError(f"unknown test {test!r}")
        U = np.zeros((4, g.ny, g.nx))
        self.lib.orc_setup(test.encode(), n, C.byref(g), C.byref(t_end), _p(U))
        return g, t_end.value, U

    def rhs(self, g, scheme, lim, U, threads=0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        L = np.empty_like(U)
        self.lib.orc_rhs(C.byref(g), SCHEME[scheme], LIMITER[lim], _p(U), _p(L), threads)
        return L

    def compute_dt(self, g, cfl, U, threads=0):
        """Returns (dt, sx, sy, bad) where bad is None or (i, j) of the first non-physical cell."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        dt, sx, sy = C.c_double(), C.c_double(), C.c_double()
        bi, bj = C.c_int(-1), C.c_int(-1)
        rc = self.lib.orc_compute_dt(C.byref(g), cfl, _p(U), C.byref(dt), C.byref(sx), C.byref(sy),
                                     C.byref(bi), C.byref(bj), threads)
        return (None if rc else dt.value), sx.value, sy.value, ((bi.value, bj.value) if rc else None)

    def step(self, g, scheme, lim, U, dt, threads=0):
        """One SSP-RK2 step, in place on a contiguous float64 (4, ny, nx) array."""
        assert U.flags.c_contiguous and U.dtype == np.float64
        self.lib.orc_step(C.byref(g), SCHEME[scheme], LIMITER[lim], _p(U), dt, threads)
        return U

    def block_average(self, g, U, m):
        U = np.ascontiguousarray(U, dtype=np.float64)
        f = g.nx // m
        out = np.empty((4, g.ny // f, m), dtype=np.float32)
        self.lib.orc_block_average(C.byref(g), _p(U), m, _p(out, _F))
        return out

    def run(self, g, scheme, lim, cfl, U, t_end, t_start=0.0, nout=1, sizes=(), max_steps=10**7,
            record_dts=False, threads=0):
        """The driver's run loop.  U is advanced in place.  Returns a dict with steps, t, dts,
        frame_t, frames {m: (nout+1, 4, ny_c, m)} and bad (None or (i, j))."""
        assert U.flags.c_contiguous and U.dtype == np.float64
        sizes = list(sizes)
        S = np.array(sizes, dtype=np.int32)
        total = sum((nout + 1) * 4 * m * (g.ny // (g.nx // m)) for m in sizes)
        frames = np.zeros(total, dtype=np.float32) if sizes else None
        frame_t = np.zeros(nout + 1)
        dts = np.zeros(max_steps) if record_dts else None
        steps, t = C.c_longlong(), C.c_double()
        bad = np.full(2, -1, dtype=np.int32)
        rc = self.lib.orc_run(C.byref(g), SCHEME[scheme], LIMITER[lim], cfl, _p(U), t_start, t_end, nout,
                              _p(S, _I) if sizes else None, len(sizes), _p(frames, _F) if sizes else None,
                              _p(frame_t), _p(dts), max_steps, C.byref(steps), C.byref(t), _p(bad, _I), threads)
        out = {"rc": rc, "steps": steps.value, "t": t.value, "frame_t": frame_t,
               "dts": dts[:steps.value] if record_dts else None,
               "bad": (int(bad[0]), int(bad[1])) if rc == 1 else None, "frames": {}}
        off = 0
        for m in sizes:
            nyc = g.ny // (g.nx // m)
            n = (nout + 1) * 4 * nyc * m
            out["frames"][m] = frames[off:off + n].reshape(nout + 1, 4, nyc, m)
            off += n
        return out


class Reference(_Lib):
    """The reference solver's own headers, built from source by oracle/Makefile (container only
    unless the prebuilt library travelled with the tree)."""

    def __init__(self, path: Path = REF_SO):
        super().__init__(path, "ref")
        L = self.lib
        L.ref_compute_dt.argtypes = [C.POINTER(Grid), C.c_double, _D, _D, _I, _I]
        L.ref_compute_dt.restype = C.c_int
        L.ref_step.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, _D, C.c_double]
        L.ref_run.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, C.c_double, _D, C.c_double, C.c_double,
                              C.c_int, _D, _D, _D, C.c_longlong, _L, _D, _I]
        L.ref_run.restype = C.c_int
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_set_threads.restype = C.c_int

    def compute_dt(self, g, cfl, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        dt = C.c_double()
        bi, bj = C.c_int(-1), C.c_int(-1)
        rc = self.lib.ref_compute_dt(C.byref(g), cfl, _p(U), C.byref(dt), C.byref(bi), C.byref(bj))
        return (None if rc else dt.value), ((bi.value, bj.value) if rc else None)

    def step(self, g, scheme, lim, U, dt):
        assert U.flags.c_contiguous and U.dtype == np.float64
        self.lib.ref_step(C.byref(g), SCHEME[scheme], LIMITER[lim], _p(U), dt)
        return U

    def run(self, g, scheme, lim, cfl, U, t_end, t_start=0.0, nout=1, keep_states=False,
            max_steps=10**7, record_dts=False, threads=0):
        """threads > 0: the reference's OpenMP loops use that many threads for this call."""
        if threads > 0:
            old = self.lib.ref_set_threads(threads)
            try:
                return self.run(g, scheme, lim, cfl, U, t_end, t_start, nout, keep_states, max_steps, record_dts)
            finally:
                self.lib.ref_set_threads(old)
        assert U.flags.c_contiguous and U.dtype == np.float64
        states = np.zeros((nout + 1,) + U.shape) if keep_states else None
        frame_t = np.zeros(nout + 1)
        dts = np.zeros(max_steps) if record_dts else None
        steps, t = C.c_longlong(), C.c_double()
        bad = np.full(2, -1, dtype=np.int32)
        rc = self.lib.ref_run(C.byref(g), SCHEME[scheme], LIMITER[lim], cfl, _p(U), t_start, t_end, nout,
                              _p(states), _p(frame_t), _p(dts), max_steps, C.byref(steps), C.byref(t),
                              _p(bad, _I))
        return {"rc": rc, "steps": steps.value, "t": t.value, "frame_t": frame_t, "states": states,
                "dts": dts[:steps.value] if record_dts else None,
                "bad": (int(bad[0]), int(bad[1])) if rc == 1 else None}


def have_reference() -> bool:
    return REF_SO.exists()


def omp_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0)

"""Binary64 building blocks on the GPU: the kernels' shared-reciprocal division and the compiler's
division and square root must give the correctly rounded IEEE results, the ones the host (and the
reference solver, compiled without contraction) computes.  Checked on the device the tests run on."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(lib, x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    qf, qi, rt = (np.empty_like(x) for _ in range(3))
    D = C.POINTER(C.c_double)
    st = lib.npde_selftest_arith(x.ctypes.data_as(D), y.ctypes.data_as(D), x.size, qf.ctypes.data_as(D),
                                 qi.ctypes.data_as(D), rt.ctypes.data_as(D))
    assert st == 0
    return qf, qi, rt


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.fixture(scope="module")
def lib(npde_mod):
    import npde._lib as L
    return L.load()


@pytest.mark.parametrize("lo,hi", [(-40, 40), (-300, 300)])
def test_division_and_sqrt_are_correctly_rounded(lib, lo, hi):
    rng = np.random.default_rng(abs(lo) * 7 + hi)
    n = 2_000_000
    sx = rng.choice([-1.0, 1.0], n)
    x = sx * np.exp2(rng.uniform(lo, hi, n)) * rng.uniform(1, 2, n)
    y = rng.choice([-1.0, 1.0], n) * np.exp2(rng.uniform(lo, hi, n)) * rng.uniform(1, 2, n)
    x[:1000] = 0.0  # zero numerators (e.g. a momentum at rest) over positive denominators
    y[:1000] = np.abs(y[:1000])
    x[1000:2000] = -0.0
    y[1000:2000] = np.abs(y[1000:2000])
    qf, qi, rt = run(lib, x, y)
    host = x / y
    assert np.array_equal(bits(qi), bits(host)), "compiler division differs from IEEE"
    assert np.array_equal(bits(qf), bits(host)), "shared-reciprocal division differs from IEEE"
    ax = np.abs(x)
    assert np.array_equal(bits(rt[x > 0]), bits(np.sqrt(ax[x > 0]))), "sqrt differs from IEEE"


def test_values_near_one_and_hard_cases(lib):
    rng = np.random.default_rng(1)
    n = 1_000_000
    # quotients close to representable midpoints: y = 1 + k ulp, x = multiples; and mantissas of all ones
    y = np.concatenate([1.0 + rng.integers(0, 2**20, n // 2) * 2.0**-52, np.full(n // 4, 2.0 - 2.0**-52),
                        rng.uniform(0.1, 10, n // 4)])
    x = np.concatenate([rng.uniform(0.5, 4, n // 2), rng.uniform(1, 2, n // 4),
                        np.nextafter(rng.uniform(0.1, 10, n // 4), np.inf)])
    qf, qi, _ = run(lib, x, y)
    assert np.array_equal(bits(qf), bits(x / y)) and np.array_equal(bits(qi), bits(x / y))


def test_signed_zeros_and_constants(lib):
    """Every denominator of the scheme is positive for a physical state (rho, gamma - 1); there the
    shared-reciprocal quotient keeps IEEE's signed zeros.  (With y < 0 a zero quotient comes out +0,
    which only states the solver already reports as non-physical can produce.)"""
    x = np.array([0.0, -0.0, 1.0, -1.0, 0.3, 1.4 * 0.3, 0.0, -0.0])
    y = np.array([1.0, 1.0, 1.4 - 1.0, 1.4 - 1.0, 0.7, 0.9, 1e-300, 3e250])
    qf, qi, _ = run(lib, x, y)
    assert np.array_equal(bits(qf), bits(x / y))
    assert np.array_equal(bits(qi), bits(x / y))
    qf, qi, _ = run(lib, np.array([0.0, -0.0]), np.array([-2.0, -2.0]))
    assert np.array_equal(bits(qi), bits(np.array([-0.0, 0.0])))  # the compiler's division is IEEE here too

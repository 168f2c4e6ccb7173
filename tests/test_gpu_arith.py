"""The device quotient and square root of the face phases against IEEE
division and square root on the host, bitwise.

The Riemann solvers, the conversions and the cfl terms divide with the steps
of the general double division sequence without its range scaling
(pip_phys.h qrcp/quo/qdiv) and take square roots with the general sequence
without the scaling of tiny arguments (qsqrt).  Both are the IEEE result on
the operand ranges stated there; pip_check_arith runs them elementwise.
"""
import ctypes

import numpy as np
import pytest

from pip_amd import _lib

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _device(n, d):
    n = np.ascontiguousarray(n, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    q, r = np.empty_like(n), np.empty_like(n)
    _lib.check(_lib.lib().pip_check_arith(_p(n), _p(d), _p(q), _p(r), n.size))
    return q, r


def _bitwise(got, want):
    """Indices where got and want differ in their bits (any NaN matches any NaN)."""
    nan = np.isnan(got) & np.isnan(want)
    return np.flatnonzero(~nan & (got.view(np.uint64) != want.view(np.uint64)))


def _doubles(rng, size, emin, emax):
    m = rng.uniform(1.0, 2.0, size)
    e = rng.integers(emin, emax + 1, size)
    s = rng.choice([-1.0, 1.0], size)
    return s * np.ldexp(m, e)


def _check_div(n, d):
    q, _ = _device(n, d)
    with np.errstate(all="ignore"):
        want = n / d
    bad = _bitwise(q, want)
    assert bad.size == 0, f"{bad.size} quotients differ, first n={n[bad[0]]!r} d={d[bad[0]]!r}"


def test_quotients_of_random_operands():
    rng = np.random.default_rng(2024)
    size = 1 << 22
    _check_div(_doubles(rng, size, -380, 380), _doubles(rng, size, -380, 380))


def test_quotients_near_the_ends_of_the_stated_range():
    # |n| >= 2^-969, |d| and 1/|d| normal, quotient normal and below 2^767
    rng = np.random.default_rng(7)
    size = 1 << 22
    en = rng.integers(-960, 1001, size)
    eq = rng.integers(-1000, 761, size)
    ed = en - eq
    keep = (ed >= -1020) & (ed <= 1020)
    n = rng.choice([-1.0, 1.0], size) * np.ldexp(rng.uniform(1.0, 2.0, size), en)
    d = rng.choice([-1.0, 1.0], size) * np.ldexp(rng.uniform(1.0, 2.0, size), ed)
    _check_div(n[keep], d[keep])


def test_quotients_close_to_rounding_ties():
    # n = fl(d q): the exact quotient lies within an ulp of q, often near a
    # midpoint of two doubles; plus exact quotients and neighbouring operands
    rng = np.random.default_rng(11)
    size = 1 << 21
    d = _doubles(rng, size, -60, 60)
    q = _doubles(rng, size, -60, 60)
    n = d * q
    ints = rng.integers(1, 1 << 20, size).astype(np.float64)
    n2 = ints * d
    n3 = np.nextafter(d, np.inf)
    n4 = np.nextafter(n, -np.inf)
    _check_div(np.concatenate([n, n2, n3, n4, d]), np.concatenate([d, d, d, d, d]))


def test_quotients_of_special_operands():
    vals = np.array([0.0, -0.0, 1.0, -1.0, 3.0, -7.5, 1e-300, 1e300, np.inf, -np.inf, np.nan,
                     2.0 ** -1000, 5e-324])
    n, d = np.meshgrid(vals, vals)
    q, _ = _device(n.ravel(), d.ravel())
    with np.errstate(all="ignore"):
        want = n.ravel() / d.ravel()
    # zeros, infinities and NaNs are exact whatever the other operand
    special = ~np.isfinite(n.ravel()) | ~np.isfinite(d.ravel()) | (n.ravel() == 0) | (d.ravel() == 0)
    bad = _bitwise(q[special], want[special])
    assert bad.size == 0, (n.ravel()[special][bad], d.ravel()[special][bad])


def test_square_roots():
    rng = np.random.default_rng(5)
    size = 1 << 22
    x = np.abs(_doubles(rng, size, -767, 1023))
    x = np.concatenate([x, np.nextafter(x, np.inf), np.square(_doubles(rng, 1 << 16, -300, 300)),
                        [0.0, -0.0, np.inf, -np.inf, np.nan, -1.0, 1.0, 4.0, 2.0 ** -767]])
    _, r = _device(x, np.ones_like(x))
    with np.errstate(all="ignore"):
        want = np.sqrt(x)
    bad = _bitwise(r, want)
    assert bad.size == 0, f"{bad.size} roots differ, first x={x[bad[0]]!r}"

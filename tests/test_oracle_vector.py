"""Pins of the FP64 vector SHT oracle (oracle/vector.py) by calculus and
invariants independent of its formulas: the two table rows are the
theta-derivative and the m / sin(theta) multiple of the CS-phase Legendre
functions; the inverse of a spheroidal (toroidal) coefficient set is the
gradient (curl) of the scalar field those coefficients synthesise; the
forward transform inverts the inverse on band-limited fields."""

import numpy as np
import pytest

from oracle.geometry import RingGrid
from oracle.legendre import legpoly
from oracle.vector import VectorSHTOracle, dlegpoly, scalar_field_at

L = 14


def _thetas():
    return np.linspace(0.15, np.pi - 0.15, 9)


def test_first_row_is_theta_derivative_of_cs_legendre():
    th = _thetas()
    h = 1e-5
    d = dlegpoly(L, L, np.cos(th))
    cs = ((-1.0) ** np.arange(L))[:, None, None]
    dP = cs * (legpoly(L, L, np.cos(th + h)) - legpoly(L, L, np.cos(th - h))) / (2 * h)
    assert np.max(np.abs(d[0] - dP)) < 1e-7 * np.max(np.abs(dP))


def test_second_row_is_m_p_over_sin_theta():
    th = _thetas()
    d = dlegpoly(L, L, np.cos(th))
    m = np.arange(L)[:, None, None]
    cs = ((-1.0) ** np.arange(L))[:, None, None]
    ref = m * cs * legpoly(L, L, np.cos(th)) / np.sin(th)
    np.testing.assert_allclose(d[1], ref, rtol=0, atol=1e-12 * np.max(np.abs(ref)))


def test_csphase_false_flips_odd_m_only():
    x = np.cos(_thetas())
    a, b = dlegpoly(L, L, x, True), dlegpoly(L, L, x, False)
    np.testing.assert_array_equal(a[:, 0::2], b[:, 0::2])
    np.testing.assert_array_equal(a[:, 1::2], -b[:, 1::2])


def _coeffs(rng, lmax, mmax):
    a = rng.standard_normal((lmax, mmax)) + 1j * rng.standard_normal((lmax, mmax))
    a[:, 0] = a[:, 0].real
    a[np.arange(lmax)[:, None] < np.arange(mmax)[None, :]] = 0
    return a


@pytest.mark.parametrize("part", ["spheroidal", "toroidal"])
def test_inverse_is_gradient_or_curl_of_scalar_synthesis(part):
    nside, lmax, mmax = 8, 7, 6
    rng = np.random.default_rng(21)
    a = _coeffs(rng, lmax, mmax)
    o = VectorSHTOracle.healpix(nside, lmax, mmax)
    c = np.zeros((2, lmax, mmax), dtype=np.complex128)
    c[0 if part == "spheroidal" else 1] = a
    v = o.inverse(c)
    g = RingGrid.healpix(nside)
    h = 1e-6
    checked = 0
    for t in range(g.nring):
        n = int(g.nphi[t])
        if n // 2 < mmax:  # ring truncation / Nyquist rules differ from the smooth field
            continue
        th = np.full(n, g.theta[t])
        ph = g.phi0[t] + 2 * np.pi * np.arange(n) / n
        dth = (scalar_field_at(a, th + h, ph) - scalar_field_at(a, th - h, ph)) / (2 * h)
        dph = (scalar_field_at(a, th, ph + h) - scalar_field_at(a, th, ph - h)) / (2 * h) / np.sin(th)
        want = (dth, dph) if part == "spheroidal" else (dph, -dth)
        got = v[:, g.offset[t]:g.offset[t] + n]
        scale = np.max(np.abs(want))
        np.testing.assert_allclose(got[0], want[0], rtol=0, atol=1e-6 * scale)
        np.testing.assert_allclose(got[1], want[1], rtol=0, atol=1e-6 * scale)
        checked += 1
    assert checked > g.nring // 2


def test_forward_inverts_inverse_on_band_limited_fields():
    nside, lmax, mmax = 32, 10, 10
    rng = np.random.default_rng(22)
    c = np.stack([_coeffs(rng, lmax, mmax), _coeffs(rng, lmax, mmax)])
    c[:, 0, :] = 0  # l = 0 has no gradient or curl
    o = VectorSHTOracle.healpix(nside, lmax, mmax)
    back = o.forward(o.inverse(c))
    err = np.linalg.norm(back - c) / np.linalg.norm(c)
    assert err < 1e-4, err
    # the two parts do not leak into each other
    only_s = c.copy()
    only_s[1] = 0
    leak = o.forward(o.inverse(only_s))[1]
    assert np.linalg.norm(leak) < 1e-4 * np.linalg.norm(c[0])


@pytest.mark.parametrize("inverse", [False, True])
def test_norms_against_scipy_unnormalised_legendre(inverse):
    """norm="schmidt" rows are built from sqrt((l-m)!/(l+m)!) P^m_l (P from
    scipy's lpmv, which carries the CS sign): the theta-derivative of that
    function at degree l, and m / sin(theta) times it with the degree-(l+1)
    factor sqrt((2l+1)/(2l+3)) the reference's ladder identity leaves; any
    other norm but "ortho" is sqrt(4 pi) times the orthonormal rows; inverse
    tables take the reciprocal factors."""
    from math import factorial
    import scipy.special as sp
    th = _thetas()
    h = 1e-6
    lmax = mmax = 9
    d_s = dlegpoly(mmax, lmax, np.cos(th), True, "schmidt", inverse)
    d_o = dlegpoly(mmax, lmax, np.cos(th), True, "ortho")
    d_4 = dlegpoly(mmax, lmax, np.cos(th), True, "four-pi", inverse)
    c4 = np.sqrt(4 * np.pi)
    np.testing.assert_allclose(d_4, d_o / c4 if inverse else d_o * c4, rtol=1e-13, atol=0)
    for l in range(lmax):
        for m in range(min(l + 1, mmax)):
            q = np.sqrt(factorial(l - m) / factorial(l + m))
            fs = (lambda t: q * sp.lpmv(m, l, np.cos(t)))
            d0 = (fs(th + h) - fs(th - h)) / (2 * h)
            d1 = np.sqrt((2 * l + 1) / (2 * l + 3)) * m * fs(th) / np.sin(th)
            if inverse:  # schmidt inverse = schmidt forward * (2L+1) / (4 pi) at the row's degree
                d0 = d0 * (2 * l + 1) / (4 * np.pi)
                d1 = d1 * (2 * l + 3) / (4 * np.pi)
            scale = max(np.max(np.abs(d0)), 1e-30)
            np.testing.assert_allclose(d_s[0, m, l], d0, rtol=0, atol=1e-7 * scale, err_msg=f"D0 l={l} m={m}")
            np.testing.assert_allclose(d_s[1, m, l], d1, rtol=0, atol=1e-12 * max(np.max(np.abs(d1)), 1.0),
                                       err_msg=f"D1 l={l} m={m}")

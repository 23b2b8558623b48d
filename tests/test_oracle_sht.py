"""Pins the FP64 SHT oracle (oracle/geometry.py, legendre.py, sht.py,
cpu_path.py) without a GPU, against closed forms, scipy's spherical
harmonics, brute-force sums and the HEALPix ring-weight tables.
"""

import os

import numpy as np
import pytest
import scipy.special as sp
import torch

from oracle import geometry as G
from oracle import legendre as L
from oracle import sht as S
from oracle.cpu_path import CPUTorchSHT

REF_DATA = "/root/reference/cuhpx/data"


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16])
def test_ring_layout(nside):
    nring = 4 * nside - 1
    nphi = [G.nphi_ring(t, nside) for t in range(nring)]
    assert sum(nphi) == 12 * nside * nside
    off = np.concatenate([[0], np.cumsum(nphi)[:-1]])
    assert [G.cumulative_nphi_ring(t, nside) for t in range(nring)] == list(off)
    # mirror symmetry of lengths and phases
    for t in range(nring):
        assert nphi[t] == nphi[nring - 1 - t]
        assert G.p2phi_ring(t, 0, nside) == pytest.approx(G.p2phi_ring(nring - 1 - t, 0, nside))


def test_ring_geometry_matches_pixel_centres():
    """z of HEALPix ring centres (Gorski et al. 2005, eqs. 4-5) and the phase of
    the first pixel of each ring (eqs. 6-8)."""
    nside = 4
    z, _ = G.healpix_weights(nside, "none")
    z = np.flip(z)  # north to south
    for t in range(4 * nside - 1):
        i = t + 1
        if i < nside:
            zt = 1 - i * i / (3 * nside * nside)
            phi0 = np.pi / (2 * i) * 0.5
        elif i <= 3 * nside:
            zt = 4 / 3 - 2 * i / (3 * nside)
            s = (i - nside + 1) % 2
            phi0 = np.pi / (2 * nside) * (0.5 * s)
        else:
            j = 4 * nside - i
            zt = -(1 - j * j / (3 * nside * nside))
            phi0 = np.pi / (2 * j) * 0.5
        assert z[t] == pytest.approx(zt, abs=1e-15)
        assert G.p2phi_ring(t, 0, nside) == pytest.approx(phi0, abs=1e-15)


@pytest.mark.parametrize("nside", [2, 16, 256, 1024])
def test_ring_weights_integrate_constant(nside):
    g = G.RingGrid.healpix(nside)
    assert (g.weight * g.nphi).sum() == pytest.approx(4 * np.pi, rel=1e-13)


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference FITS tables not present")
def test_ring_weights_match_fits_tables():
    """The committed npz equals the FITS column (re-parsed), and the column's
    extremes equal the MAXVAL1 / MINVAL1 header keywords of each table."""
    import re
    from cuhpx_amd.data.convert_ring_weights import read_fits_column0
    for name in sorted(os.listdir(REF_DATA)):
        nside = int(re.findall(r"\d+", name)[0])
        path = os.path.join(REF_DATA, name)
        w = read_fits_column0(path)
        np.testing.assert_array_equal(w, G.read_ring_weight(nside))
        head = open(path, "rb").read(8640 + 11 * 2880).decode("latin1")
        mx = float(re.search(r"MAXVAL1 =\s*([-+.\dE]+)", head).group(1))
        mn = float(re.search(r"MINVAL1 =\s*([-+.\dE]+)", head).group(1))
        assert w.max() == pytest.approx(mx, rel=1e-11)
        assert w.min() == pytest.approx(mn, rel=1e-11)


# ---------------------------------------------------------------- Legendre
def test_legendre_closed_forms():
    th = np.linspace(0.1, 3.0, 7)
    x = np.cos(th)
    p = L.legpoly(3, 3, x)
    np.testing.assert_allclose(p[0, 0], 1 / np.sqrt(4 * np.pi))
    np.testing.assert_allclose(p[0, 1], np.sqrt(3 / (4 * np.pi)) * x)
    np.testing.assert_allclose(p[1, 1], np.sqrt(3 / (8 * np.pi)) * np.sin(th))  # no CS phase
    np.testing.assert_allclose(p[0, 2], np.sqrt(5 / (16 * np.pi)) * (3 * x * x - 1))
    np.testing.assert_allclose(p[2, 2], np.sqrt(15 / (32 * np.pi)) * np.sin(th) ** 2)
    assert np.all(p[1, 0] == 0) and np.all(p[2, 1] == 0)


def test_legendre_matches_scipy():
    th = np.linspace(0.05, 3.09, 23)
    lmax = 40
    p = L.legpoly(lmax, lmax, np.cos(th))
    for l in range(lmax):
        for m in range(l + 1):
            y = sp.sph_harm_y(l, m, th, 0.0).real * (-1) ** m  # undo Condon-Shortley
            np.testing.assert_allclose(p[m, l], y, rtol=1e-10, atol=1e-13)


def test_legendre_orthonormal_on_gauss_nodes():
    x, w = np.polynomial.legendre.leggauss(64)
    p = L.legpoly(20, 40, x)
    for m in range(20):
        gram = 2 * np.pi * (p[m] * w) @ p[m].T
        np.testing.assert_allclose(gram[m:, m:], np.eye(40 - m), atol=1e-12)


def test_scaled_recurrence_equals_reference_recurrence():
    g = G.RingGrid.healpix(32)
    x = np.cos(g.theta)
    np.testing.assert_allclose(L.legpoly_scaled(65, 65, x), L.legpoly(65, 65, x), rtol=1e-12, atol=1e-13)


def test_scaled_recurrence_keeps_tiny_values():
    """Near the pole at high m the plain recurrence runs into the FP64 range
    limit; the scaled one agrees with it wherever the value is above 2^-600 and
    emits zero below (far under anything FP32 can hold)."""
    x = np.array([np.cos(0.002), np.cos(0.01), 0.3])
    a = L.legpoly(400, 420, x)
    b = L.legpoly_scaled(400, 420, x)
    ok = np.abs(a) > 2.0 ** -590
    np.testing.assert_allclose(b[ok], a[ok], rtol=1e-10)
    assert np.all(np.abs(a[b == 0]) < 2.0 ** -599)
    # values the plain recurrence flushes to zero stay zero in the scaled one
    # only because they are tiny, not because of the flush
    assert np.all(np.abs(b[a == 0]) < 2.0 ** -599)


# ---------------------------------------------------------------- ring DFT and SHT

def test_streamed_oracle_equals_table_oracle():
    """The per-m streamed tables (for grids whose full table does not fit in
    memory) give the same transforms as the full-table oracle."""
    o = S.SHTOracle.healpix(8, 20, 17)
    o_s = S.SHTOracle.healpix(8, 20, 17, streamed=True)
    assert o_s.pct is None
    rng = np.random.default_rng(0)
    f = rng.standard_normal((2, o.grid.npix))
    a = o.forward(f)
    np.testing.assert_allclose(o_s.forward(f), a, rtol=0, atol=1e-13 * np.abs(a).max())
    g = o.inverse(a)
    np.testing.assert_allclose(o_s.inverse(a), g, rtol=0, atol=1e-13 * np.abs(g).max())

@pytest.mark.parametrize("nside,mmax", [(2, 5), (4, 9), (4, 20), (8, 17)])
def test_ring_dft_matches_bruteforce(nside, mmax):
    g = G.RingGrid.healpix(nside)
    rng = np.random.default_rng(0)
    f = rng.standard_normal((2, g.npix))
    np.testing.assert_allclose(S.ring_rfft(f, g, mmax), S.dft_forward_bruteforce(f, g, mmax), atol=1e-12)
    spec = rng.standard_normal((2, g.nring, mmax)) + 1j * rng.standard_normal((2, g.nring, mmax))
    np.testing.assert_allclose(S.ring_irfft(spec, g, mmax), S.dft_inverse_bruteforce(spec, g, mmax), atol=1e-12)


def _pixel_angles(g):
    theta = np.repeat(g.theta, g.nphi)
    phi = np.concatenate([g.phi0[t] + 2 * np.pi * np.arange(g.nphi[t]) / g.nphi[t] for t in range(g.nring)])
    return theta, phi


@pytest.mark.parametrize("nside", [2, 4])
def test_forward_is_weighted_sum_against_scipy_harmonics(nside):
    """a_lm = sum_p w_p f_p P_lm(theta_p) e^{-i m phi_p} over the rings where
    m < N_t/2 + 1 (the reference's truncation), with P_lm from scipy."""
    lmax = mmax = 2 * nside + 2
    o = S.SHTOracle.healpix(nside, lmax, mmax)
    g = o.grid
    f = np.random.default_rng(1).standard_normal(g.npix)
    a = o.forward(f)
    theta, phi = _pixel_angles(g)
    wpix = np.repeat(g.weight, g.nphi)
    ringlen = np.repeat(g.nphi, g.nphi)
    for l in range(lmax):
        for m in range(min(l, mmax - 1) + 1):
            y = sp.sph_harm_y(l, m, theta, 0.0).real * (-1) ** m
            keep = m < ringlen // 2 + 1
            ref = np.sum((wpix * f * y * np.exp(-1j * m * phi))[keep])
            assert a[l, m] == pytest.approx(ref, abs=1e-12)


@pytest.mark.parametrize("nside", [2, 4])
def test_inverse_of_unit_coefficient_is_scipy_harmonic(nside):
    lmax = mmax = 2 * nside + 1
    o = S.SHTOracle.healpix(nside, lmax, mmax)
    g = o.grid
    theta, phi = _pixel_angles(g)
    ringlen = np.repeat(g.nphi, g.nphi)
    for l, m in [(0, 0), (3, 0), (3, 2), (lmax - 1, mmax - 1), (lmax - 1, 3)]:
        a = np.zeros((lmax, mmax), dtype=np.complex128)
        a[l, m] = 0.7 - 0.2j
        f = o.inverse(a)
        y = sp.sph_harm_y(l, m, theta, 0.0).real * (-1) ** m
        c = np.where(m == 0, 1.0, np.where(2 * m == ringlen, 1.0, 2.0))
        ref = c * np.real(a[l, m] * y * np.exp(1j * m * phi))
        ref = np.where(m < ringlen // 2 + 1, ref, 0.0)
        np.testing.assert_allclose(f, ref, atol=1e-12)


def test_band_limited_round_trip_error_is_small():
    """HEALPix ring quadrature is not exact; for lmax = nside + 1 the analysis
    of a band-limited map recovers its coefficients to ~1e-4, and the error
    falls as nside grows at fixed lmax / nside."""
    errs = []
    for nside in (16, 32):
        lmax = mmax = nside + 1
        o = S.SHTOracle.healpix(nside, lmax, mmax)
        rng = np.random.default_rng(2)
        a = (rng.standard_normal((lmax, mmax)) + 1j * rng.standard_normal((lmax, mmax)))
        a[:, 0] = a[:, 0].real
        a = np.tril(a)
        back = o.forward(o.inverse(a))
        errs.append(np.linalg.norm(back - a) / np.linalg.norm(a))
    assert errs[0] < 2e-4 and errs[1] < errs[0]


@pytest.mark.parametrize("nside,lmax,mmax", [(2, 6, 6), (4, 9, 9), (4, 15, 12)])
def test_exact_adjoints_dot_product(nside, lmax, mmax):
    o = S.SHTOracle.healpix(nside, lmax, mmax)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(o.grid.npix)
    a = rng.standard_normal((lmax, mmax)) + 1j * rng.standard_normal((lmax, mmax))
    # real pairing <u, v> = Re sum conj(u) v
    lhs = np.real(np.vdot(a, o.forward(f)))
    rhs = np.dot(f, o.forward_adjoint(a))
    assert lhs == pytest.approx(rhs, rel=1e-11)
    g = rng.standard_normal(o.grid.npix)
    lhs = np.dot(g, o.inverse(a))
    rhs = np.real(np.vdot(o.inverse_adjoint(g), a))
    assert lhs == pytest.approx(rhs, rel=1e-11)


def test_exact_adjoints_match_torch_autograd():
    """torch autograd through torch.fft on the same operator (FP64)."""
    nside, lmax, mmax = 4, 9, 9
    o = S.SHTOracle.healpix(nside, lmax, mmax)
    g = o.grid
    pct = torch.from_numpy(o.pct)
    w = torch.from_numpy(g.weight)

    def fwd(f):
        rows = []
        for t in range(g.nring):
            n, off = int(g.nphi[t]), int(g.offset[t])
            k = min(mmax, n // 2 + 1)
            spec = torch.fft.rfft(f[off:off + n])[:k]
            spec = torch.cat([spec, torch.zeros(mmax - k, dtype=spec.dtype)])
            rows.append(spec * torch.exp(-1j * torch.arange(mmax, dtype=torch.float64) * float(g.phi0[t])) * w[t])
        ftm = torch.stack(rows)
        return torch.einsum("km,mlk->lm", ftm, pct.to(ftm.dtype))

    f = torch.randn(g.npix, dtype=torch.float64, requires_grad=True)
    ga = torch.randn(lmax, mmax, dtype=torch.complex128)
    out = fwd(f)
    (out.real * ga.real + out.imag * ga.imag).sum().backward()
    np.testing.assert_allclose(f.grad.numpy(), o.forward_adjoint(ga.numpy()), atol=1e-12)


def test_compat_adjoints_differ_only_on_dc_and_nyquist():
    nside, lmax, mmax = 4, 9, 9
    o = S.SHTOracle.healpix(nside, lmax, mmax)
    a = np.zeros((lmax, mmax), dtype=np.complex128)
    a[5, 3] = 1.0 + 0.5j
    np.testing.assert_allclose(o.forward_adjoint_compat(a), o.forward_adjoint(a), atol=1e-14)
    a0 = np.zeros_like(a)
    a0[5, 0] = 1.0
    np.testing.assert_allclose(o.forward_adjoint_compat(a0), 0.5 * o.forward_adjoint(a0), atol=1e-14)
    gf = np.random.default_rng(4).standard_normal(o.grid.npix)
    ex, co = o.inverse_adjoint(gf), o.inverse_adjoint_compat(gf)
    np.testing.assert_allclose(co[:, 0], 2 * ex[:, 0], atol=1e-12)
    np.testing.assert_allclose(co[:, 3], ex[:, 3], atol=1e-12)


def test_cpu_torch_path_restatement_matches_oracle():
    nside = 8
    lmax = mmax = 2 * nside + 1
    o = S.SHTOracle.healpix(nside, lmax, mmax)
    c = CPUTorchSHT(nside, lmax, mmax)
    f = np.random.default_rng(5).standard_normal(o.grid.npix)
    a = c.forward(torch.from_numpy(f)).numpy()
    # the reference builds its phase factors in complex64 even for FP64 maps
    # (hpx_sht.py:52, 67: -1j * arange(L) promotes to complex64)
    np.testing.assert_allclose(a, o.forward(f), rtol=0, atol=1e-7 * np.abs(a).max())
    back = c.inverse(torch.from_numpy(a)).numpy()
    np.testing.assert_allclose(back, o.inverse(a), rtol=0, atol=1e-7 * np.abs(back).max())


def test_l_outer_scaled_recurrence_equals_per_m_recurrence():
    """legpoly_scaled_lrows (all m advanced together, l outer) reproduces
    legpoly_scaled bit for bit, deep into the rescaled (underflow) region."""
    x = np.cos(np.linspace(1e-4, np.pi - 1e-4, 301))
    for mmax, lmax in [(9, 9), (5, 12), (12, 5), (700, 900)]:
        ref = L.legpoly_scaled(mmax, lmax, x)
        got = np.zeros_like(ref)
        for l, q in L.legpoly_scaled_lrows(mmax, lmax, x):
            assert q.shape[0] == min(l + 1, mmax)
            got[:q.shape[0], l, :] = q
        assert np.array_equal(got, ref), (mmax, lmax)


def test_streamed_forward_and_inverse_share_one_sweep():
    """forward_and_inverse equals forward and inverse called one by one."""
    o = S.SHTOracle.healpix(8, 17, 17, streamed=True)
    rng = np.random.default_rng(12)
    f = rng.standard_normal((2, o.grid.npix))
    a = rng.standard_normal((3, 17, 17)) + 1j * rng.standard_normal((3, 17, 17))
    fa, ga = o.forward_and_inverse(f, a)
    np.testing.assert_allclose(fa, o.forward(f), rtol=0, atol=1e-13)
    np.testing.assert_allclose(ga, o.inverse(a), rtol=0, atol=1e-13)

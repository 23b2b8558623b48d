"""FP64 transforms: float64 maps give complex128 coefficients and back,
computed in FP64 end to end (the reference's dtype dispatch,
src/harmonic_transform/hpx_fft.cpp:45, with the table cast to the input's
dtype, cuhpx/hpx_sht.py:557, 599).  Compared with the FP64 oracle of the same
operator within 1e-10 relative L2 (both sides round only in FP64)."""

import numpy as np
import pytest
import torch

import cuhpx_amd
from cuhpx_amd import cuhpx_fft
from cuhpx_amd.hpx_sht import ShtPlan
from oracle import sht as S
from oracle.geometry import RingGrid

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def rel_l2(x, ref):
    return float(np.linalg.norm(np.asarray(x) - ref) / np.linalg.norm(ref))


def band_limited_alm(lmax, mmax, nfield, seed):
    rng = np.random.default_rng(seed)
    l = np.arange(lmax)[:, None]
    cl = (l + 1.0) ** -2
    a = (rng.standard_normal((nfield, lmax, mmax)) + 1j * rng.standard_normal((nfield, lmax, mmax)))
    a *= np.sqrt(cl / 2)
    a[:, :, 0] = rng.standard_normal((nfield, lmax)) * np.sqrt(cl[:, 0])
    a[:, l[:, 0][:, None] < np.arange(mmax)[None, :]] = 0
    return a


@pytest.mark.parametrize("nside,lmax,mmax", [(2, 5, 5), (4, 9, 9), (4, 15, 9), (8, 17, 17), (8, 31, 17),
                                             (16, 20, 33), (32, 65, 65), (64, 129, 129)])
def test_f64_forward_inverse_match_oracle(nside, lmax, mmax):
    o = S.SHTOracle.healpix(nside, lmax, mmax)
    rng = np.random.default_rng(41)
    f = rng.standard_normal((3, o.grid.npix))
    sht = cuhpx_amd.SHTCUDA(nside, lmax=lmax, mmax=mmax)
    isht = cuhpx_amd.iSHTCUDA(nside, lmax=lmax, mmax=mmax)
    a = sht(torch.from_numpy(f).cuda())
    assert a.dtype == torch.complex128 and a.shape == (3, lmax, mmax)
    ef = rel_l2(a.cpu().numpy(), o.forward(f))
    alm = band_limited_alm(lmax, mmax, 3, 42)
    g = isht(torch.from_numpy(alm).cuda())
    assert g.dtype == torch.float64 and g.shape == (3, o.grid.npix)
    ei = rel_l2(g.cpu().numpy(), o.inverse(alm))
    print(f"[f64 nside={nside} L={lmax} M={mmax}] forward {ef:.2e} inverse {ei:.2e}", flush=True)
    assert ef < TOL64 and ei < TOL64


def test_f64_nside256_band_limited():
    """nside 256, lmax = mmax = 513 (every FFT class of the FP32 path: 1024-,
    2048-point Bluestein and power-of-two rings), 2 band-limited fields."""
    nside, L = 256, 513
    o = S.SHTOracle.healpix(nside, L, L, streamed=True)
    alm = band_limited_alm(L, L, 2, 43)
    f64 = o.inverse(alm)
    sht = cuhpx_amd.SHTCUDA(nside, lmax=L, mmax=L)
    isht = cuhpx_amd.iSHTCUDA(nside, lmax=L, mmax=L)
    a = sht(torch.from_numpy(f64).cuda()).cpu().numpy()
    g = isht(torch.from_numpy(alm).cuda()).cpu().numpy()
    a_ref, g_ref = o.forward_and_inverse(f64, alm)
    ef, ei = rel_l2(a, a_ref), rel_l2(g, g_ref)
    print(f"[f64 nside=256 L=513] forward {ef:.2e} inverse {ei:.2e}", flush=True)
    assert ef < TOL64 and ei < TOL64


@pytest.mark.parametrize("mode", ["exact", "compat"])
def test_f64_adjoints_match_oracle(mode):
    nside, L = 8, 17
    o = S.SHTOracle.healpix(nside, L, L)
    plan = ShtPlan.healpix(nside, L, L)
    rng = np.random.default_rng(44)
    ga = rng.standard_normal((2, L, L)) + 1j * rng.standard_normal((2, L, L))
    gf = rng.standard_normal((2, o.grid.npix))
    got_f = plan.forward_adjoint(torch.from_numpy(ga).cuda(), "fp32", mode).cpu().numpy()
    got_a = plan.inverse_adjoint(torch.from_numpy(gf).cuda(), "fp32", mode).cpu().numpy()
    if mode == "exact":
        ref_f, ref_a = o.forward_adjoint(ga), o.inverse_adjoint(gf)
    else:
        ref_f, ref_a = o.forward_adjoint_compat(ga), o.inverse_adjoint_compat(gf)
    assert got_f.dtype == np.float64 and got_a.dtype == np.complex128
    assert rel_l2(got_f, ref_f) < TOL64 and rel_l2(got_a, ref_a) < TOL64


def test_f64_autograd_through_modules():
    nside, L = 8, 17
    o = S.SHTOracle.healpix(nside, L, L)
    sht = cuhpx_amd.SHTCUDA(nside, lmax=L, mmax=L)
    isht = cuhpx_amd.iSHTCUDA(nside, lmax=L, mmax=L)
    rng = np.random.default_rng(45)
    f = torch.from_numpy(rng.standard_normal(o.grid.npix)).cuda().requires_grad_(True)
    w = torch.from_numpy(rng.standard_normal((L, L))).cuda()
    a = sht(f)
    ((a.real * w).sum() + (a.imag * w).sum()).backward()
    assert f.grad.dtype == torch.float64
    assert rel_l2(f.grad.cpu().numpy(), o.forward_adjoint((w + 1j * w).cpu().numpy())) < TOL64
    a2 = torch.from_numpy(band_limited_alm(L, L, 1, 46)[0]).cuda().requires_grad_(True)
    g = torch.from_numpy(rng.standard_normal(o.grid.npix)).cuda()
    (isht(a2) * g).sum().backward()
    assert a2.grad.dtype == torch.complex128
    assert rel_l2(a2.grad.cpu().numpy(), o.inverse_adjoint(g.cpu().numpy())) < TOL64


@pytest.mark.parametrize("nside,L", [(4, 9), (16, 20), (64, 129)])
def test_f64_ring_ffts_match_oracle(nside, L):
    g = RingGrid.healpix(nside)
    rng = np.random.default_rng(47)
    f = rng.standard_normal((2, g.npix))
    ftm = cuhpx_fft.healpix_rfft_batch(torch.from_numpy(f).cuda(), L, nside)
    assert ftm.dtype == torch.complex128
    ref = S.ring_rfft(f, g, L)
    assert rel_l2(ftm.cpu().numpy(), ref) < 1e-12
    back = cuhpx_fft.healpix_irfft_batch(ftm, L, nside)
    assert back.dtype == torch.float64
    assert rel_l2(back.cpu().numpy(), S.ring_irfft(ref, g, L)) < 1e-12


def test_f64_regridding_round_trip_matches_oracle():
    """Regridding.execute on float64 fields (HEALPix -> equiangular -> HEALPix)
    in FP64, against the FP64 oracle of the same chain."""
    from cuhpx_amd import Grid, Regridding
    from oracle.equiangular import latlon_grid
    nside = 8
    L = 2 * nside + 1
    hpx2eq = Regridding(Grid("healpix", nside), Grid("equiangular", (2 * nside, 4 * nside)), lmax=L, mmax=L)
    eq2hpx = Regridding(Grid("equiangular", (2 * nside, 4 * nside)), Grid("healpix", nside), lmax=L, mmax=L)
    oh = S.SHTOracle.healpix(nside, L, L)
    oe = S.SHTOracle(latlon_grid(2 * nside, 4 * nside, "equiangular", True), L, L)
    f = np.random.default_rng(48).standard_normal((2, 12 * nside * nside))
    eq = hpx2eq.execute(torch.from_numpy(f).cuda())
    assert eq.dtype == torch.float64
    back = eq2hpx.execute(eq).cpu().numpy()
    ref_eq = oe.inverse(oh.forward(f))
    assert rel_l2(eq.cpu().numpy().reshape(2, -1), ref_eq) < TOL64
    assert rel_l2(back, oh.inverse(oe.forward(ref_eq))) < TOL64


@pytest.mark.parametrize("nside,lmax,mmax,norm", [(4, 9, 7, "ortho"), (8, 31, 17, "ortho"), (16, 33, 33, "ortho"),
                                                  (8, 17, 13, "schmidt")])
def test_f64_vector_transforms_match_oracle(nside, lmax, mmax, norm):
    """VectorSHT / VectoriSHT on float64 / complex128 compute in FP64, against
    the FP64 vector oracle."""
    from oracle.vector import VectorSHTOracle
    o = VectorSHTOracle.healpix(nside, lmax, mmax, norm=norm)
    rng = np.random.default_rng(49)
    f = rng.standard_normal((2, 2, o.grid.npix))
    fwd = cuhpx_amd.VectorSHT(nside, lmax=lmax, mmax=mmax, norm=norm)
    inv = cuhpx_amd.VectoriSHT(nside, lmax=lmax, mmax=mmax, norm=norm)
    c = fwd(torch.from_numpy(f).cuda())
    assert c.dtype == torch.complex128 and c.shape == (2, 2, lmax, mmax)
    ef = rel_l2(c.cpu().numpy(), o.forward(f))
    a = rng.standard_normal((2, 2, lmax, mmax)) + 1j * rng.standard_normal((2, 2, lmax, mmax))
    a[..., np.arange(lmax)[:, None] < np.arange(mmax)[None, :]] = 0
    a[..., 0] = a[..., 0].real
    g = inv(torch.from_numpy(a).cuda())
    assert g.dtype == torch.float64 and g.shape == (2, 2, o.grid.npix)
    ei = rel_l2(g.cpu().numpy(), o.inverse(a))
    print(f"[f64 vector nside={nside} L={lmax} M={mmax} {norm}] forward {ef:.2e} inverse {ei:.2e}", flush=True)
    assert ef < TOL64 and ei < TOL64


@pytest.mark.timeout(900)
@pytest.mark.parametrize("nside", [2048, 4096])
def test_f64_large_nside_ring_by_ring(nside):
    """FP64 at nside 2048 (polar rings of 4100..8188 pixels ring by ring as
    half-length sequences: FP64 FFTs stop at 8192 points) and nside 4096
    (quarter-length Bluestein DFTs, the 16384-pixel equator rings as one
    8192-point FFT each), lmax = mmax = 9, one field, against the streamed
    FP64 oracle."""
    L = 9
    print(f"[f64 nside={nside}] oracle", flush=True)
    o = S.SHTOracle.healpix(nside, L, L, streamed=True)
    alm = band_limited_alm(L, L, 1, 47)
    f = o.inverse(alm)
    sht = cuhpx_amd.SHTCUDA(nside, lmax=L, mmax=L)
    isht = cuhpx_amd.iSHTCUDA(nside, lmax=L, mmax=L)
    print(f"[f64 nside={nside}] GPU", flush=True)
    a = sht(torch.from_numpy(f).cuda()).cpu().numpy()
    g = isht(torch.from_numpy(alm).cuda()).cpu().numpy()
    print(f"[f64 nside={nside}] oracle check", flush=True)
    a_ref, g_ref = o.forward_and_inverse(f, alm)
    ef, ei = rel_l2(a, a_ref), rel_l2(g, g_ref)
    print(f"[f64 nside={nside}] forward {ef:.2e} inverse {ei:.2e}", flush=True)
    assert ef < TOL64 and ei < TOL64


@pytest.mark.parametrize("nphi", [6002, 12004, 16384])
def test_f64_ring_by_ring_on_a_ring_grid(nphi):
    """Five rings of nphi pixels: 6002 runs ring by ring as 3001-point
    Bluestein DFTs, 12004 as two interleaved 3001-point DFTs per ring, 16384 as
    one 8192-point FFT per ring (FP64 FFTs stop at 8192 points); forward,
    inverse and both adjoints against the oracle."""
    lmax = mmax = 12
    theta = np.array([0.3, 0.9, np.pi / 2, np.pi - 0.9, np.pi - 0.3])
    nph = np.full(5, nphi)
    phi0 = np.full(5, 0.37)
    w = np.array([0.7, 1.1, 1.3, 1.1, 0.7]) * 4 * np.pi / (nphi * 4.9)
    grid = RingGrid(theta, nph, phi0, w)
    o = S.SHTOracle(grid, lmax, mmax)
    plan = ShtPlan.rings(nph, theta, phi0, w, lmax, mmax)
    rng = np.random.default_rng(48)
    f = rng.standard_normal((2, grid.npix))
    a = rng.standard_normal((2, lmax, mmax)) + 1j * rng.standard_normal((2, lmax, mmax))
    a[:, np.arange(lmax)[:, None] < np.arange(mmax)[None, :]] = 0
    ft, at = torch.from_numpy(f).cuda(), torch.from_numpy(a).cuda()
    checks = [("forward", plan.forward(ft, "fp32"), o.forward(f)),
              ("inverse", plan.inverse(at, "fp32"), o.inverse(a)),
              ("forward adjoint", plan.forward_adjoint(at, "fp32", "exact"), o.forward_adjoint(a)),
              ("inverse adjoint", plan.inverse_adjoint(ft, "fp32", "exact"), o.inverse_adjoint(f))]
    for name, got, ref in checks:
        assert got.dtype in (torch.float64, torch.complex128), name
        err = rel_l2(got.cpu().numpy(), ref)
        print(f"[f64 ring grid nphi={nphi}] {name} rel-L2 {err:.2e}", flush=True)
        assert err < TOL64, name

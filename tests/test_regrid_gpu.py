"""Lat/lon SHT legs and Regridding vs the FP64 oracle, and the gradient of a
regridding loss vs central finite differences and the reference's formula."""

import numpy as np
import pytest
import torch

import cuhpx_amd
from cuhpx_amd import Grid, Regridding
from oracle import sht as S
from oracle.equiangular import latlon_grid

pytestmark = pytest.mark.gpu


def rel_l2(x, ref):
    return float(np.linalg.norm(np.asarray(x) - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("grid", ["equiangular", "legendre-gauss", "lobatto"])
@pytest.mark.parametrize("csphase", [True, False])
def test_latlon_legs_match_oracle(grid, csphase):
    nlat, nlon, L = 16, 32, 15
    o = S.SHTOracle(latlon_grid(nlat, nlon, grid, csphase), L, L)
    fwd = cuhpx_amd.LatLonSHT(nlat, nlon, lmax=L, mmax=L, grid=grid, csphase=csphase)
    inv = cuhpx_amd.InverseLatLonSHT(nlat, nlon, lmax=L, mmax=L, grid=grid, csphase=csphase)
    rng = np.random.default_rng(0)
    f = rng.standard_normal((2, nlat, nlon)).astype(np.float32)
    a = fwd(torch.from_numpy(f).cuda()).cpu().numpy()
    assert rel_l2(a, o.forward(f.reshape(2, -1).astype(np.float64))) < 1e-5
    back = inv(torch.from_numpy(a).cuda()).cpu().numpy()
    assert back.shape == (2, nlat, nlon)
    assert rel_l2(back.reshape(2, -1), o.inverse(a.astype(np.complex128))) < 1e-5


def test_csphase_is_minus_one_to_the_m():
    nlat, nlon, L = 16, 32, 15
    f = torch.randn(nlat, nlon, device="cuda")
    a1 = cuhpx_amd.LatLonSHT(nlat, nlon, lmax=L, mmax=L, csphase=True)(f)
    a0 = cuhpx_amd.LatLonSHT(nlat, nlon, lmax=L, mmax=L, csphase=False)(f)
    sign = torch.tensor([(-1.0) ** m for m in range(L)], device="cuda")
    torch.testing.assert_close(a1, a0 * sign, atol=1e-6, rtol=1e-5)


def test_regridding_round_trip_matches_oracle():
    nside = 8
    L = 2 * nside + 1
    hpx2eq = Regridding(Grid("healpix", nside), Grid("equiangular", (2 * nside, 4 * nside)), lmax=L, mmax=L)
    eq2hpx = Regridding(Grid("equiangular", (2 * nside, 4 * nside)), Grid("healpix", nside), lmax=L, mmax=L)
    oh = S.SHTOracle.healpix(nside, L, L)
    oe = S.SHTOracle(latlon_grid(2 * nside, 4 * nside, "equiangular", True), L, L)
    f = np.random.default_rng(1).standard_normal((2, 12 * nside * nside)).astype(np.float32)
    eq = hpx2eq.execute(torch.from_numpy(f).cuda())
    assert eq.shape == (2, 2 * nside, 4 * nside)
    back = eq2hpx.execute(eq).cpu().numpy()
    ref_eq = oe.inverse(oh.forward(f.astype(np.float64)))
    ref = oh.inverse(oe.forward(ref_eq))
    assert rel_l2(eq.cpu().numpy().reshape(2, -1), ref_eq) < 1e-5
    assert rel_l2(back, ref) < 1e-5


def test_default_band_limits_follow_reference():
    r = Regridding(Grid("healpix", 4), Grid("equiangular", (8, 16)))
    assert (r.lmax, r.mmax) == (8, 9)
    r = Regridding(Grid("equiangular", (6, 20)), Grid("healpix", 4))
    assert (r.lmax, r.mmax) == (6, 9)


@pytest.mark.parametrize("adjoint", ["exact", "compat"])
def test_regridding_gradient_nside256(adjoint):
    """Loss = mean((eq2hpx(hpx2eq(f)) - target)^2) at nside 256 <-> 512 x 1024,
    lmax = mmax = 513, over a batch of 16 fields.  The chain is linear and acts
    on each field alone, so the gradient of field i is the FP64 oracle's
    gradient of that field's term, 2 r_i / (16 npix): the GPU gradient
    (hand-written adjoints, FP32) of the first and the last field is compared
    with it; in exact mode the oracle gradient is itself checked against
    central finite differences (h = 1e-3) along random directions."""
    nside, nlat, nlon, L, B = 256, 512, 1024, 513, 16
    chk = [0, B - 1]
    print(f"[regrid-grad {adjoint} B={B}]", flush=True)
    hpx2eq = Regridding(Grid("healpix", nside), Grid("equiangular", (nlat, nlon)), lmax=L, mmax=L,
                        adjoint=adjoint)
    eq2hpx = Regridding(Grid("equiangular", (nlat, nlon)), Grid("healpix", nside), lmax=L, mmax=L,
                        adjoint=adjoint)
    oh = S.SHTOracle.healpix(nside, L, L)
    oe = S.SHTOracle(latlon_grid(nlat, nlon, "equiangular", True), L, L)
    rng = np.random.default_rng(3)
    npix = 12 * nside * nside
    f = rng.standard_normal((B, npix)).astype(np.float32)
    target = rng.standard_normal((B, npix)).astype(np.float32)
    x = torch.from_numpy(f).cuda().requires_grad_(True)
    loss = ((eq2hpx.execute(hpx2eq.execute(x)) - torch.from_numpy(target).cuda()) ** 2).mean()
    loss.backward()
    g_gpu = x.grad.cpu().numpy()[chk]

    def chain(v):
        return oh.inverse(oe.forward(oe.inverse(oh.forward(v))))

    f64, t64 = f[chk].astype(np.float64), target[chk].astype(np.float64)
    r = chain(f64) - t64
    dh = 2 * r / (B * npix)
    if adjoint == "exact":
        g_ref = oh.forward_adjoint(oe.inverse_adjoint(oe.forward_adjoint(oh.inverse_adjoint(dh))))
    else:
        g_ref = oh.forward_adjoint_compat(oe.inverse_adjoint(oe.forward_adjoint(oh.inverse_adjoint_compat(dh))))
    err = rel_l2(g_gpu, g_ref)
    print(f"[regrid-grad {adjoint}] GPU vs FP64 gradient rel-L2 {err:.2e}", flush=True)
    assert err < 1e-4
    if adjoint == "exact":
        # the checked fields' share of the batch loss
        loss_fn = lambda v: float(((chain(v) - t64) ** 2).sum()) / (B * npix)
        h = 1e-3
        for i in range(3):
            d = rng.standard_normal(f64.shape)
            d /= np.linalg.norm(d)
            fd = (loss_fn(f64 + h * d) - loss_fn(f64 - h * d)) / (2 * h)
            assert abs(fd - float((g_ref * d).sum())) <= 1e-6 * abs(fd) + 1e-12
            assert abs(fd - float((g_gpu * d).sum())) <= 1e-3 * abs(fd) + 1e-9

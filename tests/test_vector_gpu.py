"""Vector SHT kernels (VectorSHT / VectoriSHT) vs the FP64 vector oracle.

Tolerances (relative L2 against the FP64 oracle of the same discrete
operator): FP32 path 1e-5, split-BF16 (3 terms) 2e-5, single-term BF16 1e-2.
"""

import ctypes

import numpy as np
import pytest
import torch

import cuhpx_amd
from cuhpx_amd import _lib
from cuhpx_amd.hpx_sht import ShtPlan
from oracle.vector import VectorSHTOracle

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-5, "bf16x3": 2e-5, "bf16": 1e-2}


def rel_l2(x, ref):
    return float(np.linalg.norm(np.asarray(x) - ref) / np.linalg.norm(ref))


def vector_coeffs(lmax, mmax, nfield, seed):
    """(S, T) coefficients with C_l = (l+1)^-2, m = 0 real, zero for m > l."""
    rng = np.random.default_rng(seed)
    l = np.arange(lmax)[:, None]
    cl = (l + 1.0) ** -2
    c = (rng.standard_normal((nfield, 2, lmax, mmax)) + 1j * rng.standard_normal((nfield, 2, lmax, mmax)))
    c *= np.sqrt(cl / 2)
    c[..., 0] = c[..., 0].real * np.sqrt(2)
    c[..., l[:, 0][:, None] < np.arange(mmax)[None, :]] = 0
    return c


@pytest.mark.parametrize("nside,lmax,mmax", [(2, 5, 5), (4, 9, 7), (8, 31, 17), (8, 12, 17),
                                             (16, 63, 33), (32, 40, 40)])
@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16"])
def test_vector_forward_inverse_match_oracle(nside, lmax, mmax, precision):
    o = VectorSHTOracle.healpix(nside, lmax, mmax)
    rng = np.random.default_rng(31)
    f = rng.standard_normal((3, 2, o.grid.npix)).astype(np.float32)
    fwd = cuhpx_amd.VectorSHT(nside, lmax=lmax, mmax=mmax, precision=precision)
    inv = cuhpx_amd.VectoriSHT(nside, lmax=lmax, mmax=mmax, precision=precision)
    c = fwd(torch.from_numpy(f).cuda()).cpu().numpy()
    c_ref = o.forward(f.astype(np.float64))
    assert c.shape == (3, 2, lmax, mmax) and c.dtype == np.complex64
    ef = rel_l2(c, c_ref)
    a = vector_coeffs(lmax, mmax, 3, 32).astype(np.complex64)
    g = inv(torch.from_numpy(a).cuda()).cpu().numpy()
    g_ref = o.inverse(a.astype(np.complex128))
    assert g.shape == (3, 2, o.grid.npix) and g.dtype == np.float32
    ei = rel_l2(g, g_ref)
    print(f"[vector nside{nside} L{lmax} M{mmax} {precision}] forward {ef:.2e} inverse {ei:.2e}")
    assert ef < TOL[precision] and ei < TOL[precision]


@pytest.mark.parametrize("csphase", [True, False])
def test_vector_default_sizes_and_csphase(csphase):
    nside = 32
    o = VectorSHTOracle.healpix(nside, csphase=csphase)
    fwd = cuhpx_amd.VectorSHT(nside, csphase=csphase)
    inv = cuhpx_amd.VectoriSHT(nside, csphase=csphase)
    assert (fwd.lmax, fwd.mmax) == (4 * nside - 1, 2 * nside + 1)
    a = vector_coeffs(fwd.lmax, fwd.mmax, 2, 33)
    g_ref = o.inverse(a)
    g = inv(torch.from_numpy(a.astype(np.complex64)).cuda()).cpu().numpy()
    assert rel_l2(g, g_ref) < TOL["fp32"]
    c = fwd(torch.from_numpy(g_ref.astype(np.float32)).cuda()).cpu().numpy()
    assert rel_l2(c, o.forward(g_ref.astype(np.float32).astype(np.float64))) < TOL["fp32"]


def test_vector_leading_dims_and_batch_equivalence():
    nside = 8
    fwd = cuhpx_amd.VectorSHT(nside, precision="bf16x3")
    inv = cuhpx_amd.VectoriSHT(nside, precision="bf16x3")
    rng = np.random.default_rng(34)
    f = torch.from_numpy(rng.standard_normal((2, 3, 2, 12 * nside * nside)).astype(np.float32)).cuda()
    c = fwd(f)
    assert c.shape == (2, 3, 2, fwd.lmax, fwd.mmax)
    one = fwd(f[1, 2])
    assert one.shape == (2, fwd.lmax, fwd.mmax)
    # launch shapes never change the arithmetic of an element: one field
    # alone equals the same field inside a batch, bit for bit
    assert torch.equal(one, c[1, 2])
    back = inv(c)
    assert back.shape == f.shape
    assert torch.equal(inv(c[0, 1]), back[0, 1])


def test_vector_inverse_leaves_input_unchanged():
    inv = cuhpx_amd.VectoriSHT(8)
    a = torch.from_numpy(vector_coeffs(inv.lmax, inv.mmax, 2, 35).astype(np.complex64)).cuda()
    keep = a.clone()
    inv(a)
    torch.cuda.synchronize()
    assert torch.equal(a, keep)


def test_vector_round_trip_band_limited():
    nside, L = 64, 24
    fwd = cuhpx_amd.VectorSHT(nside, lmax=L, mmax=L)
    inv = cuhpx_amd.VectoriSHT(nside, lmax=L, mmax=L)
    c = vector_coeffs(L, L, 2, 36)
    c[:, :, 0, :] = 0  # l = 0 carries no vector field
    ct = torch.from_numpy(c.astype(np.complex64)).cuda()
    back = fwd(inv(ct)).cpu().numpy()
    assert rel_l2(back, c) < 1e-4


def test_vector_input_checks():
    nside = 4
    fwd = cuhpx_amd.VectorSHT(nside)
    inv = cuhpx_amd.VectoriSHT(nside)
    npix = 12 * nside * nside
    with pytest.raises(ValueError):
        fwd(torch.zeros((2, npix), dtype=torch.complex64, device="cuda"))
    with pytest.raises(TypeError):
        fwd(torch.zeros((2, npix), dtype=torch.float16, device="cuda"))
    with pytest.raises(RuntimeError):
        fwd(torch.zeros((3, npix), device="cuda"))
    with pytest.raises(ValueError):
        inv(torch.zeros((2, inv.lmax, inv.mmax), device="cuda"))
    with pytest.raises(RuntimeError):
        inv(torch.zeros((2, inv.lmax + 1, inv.mmax), dtype=torch.complex64, device="cuda"))
    with pytest.raises(ValueError):
        cuhpx_amd.VectorSHT(nside, grid="equiangular")


def test_scalar_and_vector_plans_reject_each_others_calls():
    nside = 4
    vplan = ShtPlan.healpix_vector(nside, 15, 9)
    splan = ShtPlan.healpix(nside, 15, 9)
    lib = _lib.lib()
    f = torch.zeros((1, 2, 12 * nside * nside), device="cuda")
    a = torch.zeros((1, 2, 15, 9), dtype=torch.complex64, device="cuda")
    ws, nbytes = vplan.workspace(1)
    s = _lib.stream_ptr(f.device)
    rc = lib.hpx_sht_forward(vplan.handle, 0, _lib.ptr(f), _lib.ptr(a), 1, _lib.ptr(ws), nbytes, s)
    assert rc == 1 and b"vector" in lib.hpx_last_error()
    rc = lib.hpx_vsht_forward(splan.handle, 0, _lib.ptr(f), _lib.ptr(a), 1, _lib.ptr(ws), nbytes, s)
    assert rc == 1 and b"scalar" in lib.hpx_last_error()
    rc = lib.hpx_vsht_forward(vplan.handle, 0, _lib.ptr(f), _lib.ptr(a), 1, _lib.ptr(ws), nbytes - 4, s)
    assert rc == 1 and b"workspace" in lib.hpx_last_error()
    h = ctypes.c_void_p()
    w = np.ones(4 * nside - 1)
    rc = lib.hpx_vsht_plan_create_healpix(nside, 0, 9, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, 0, 0, s,
                                          ctypes.byref(h))
    assert rc == 1


@pytest.mark.parametrize("norm", ["schmidt", "four-pi"])
def test_vector_norms_match_oracle(norm):
    """norm scales the tables as the reference's legpoly does (VectorSHT with
    the forward factors, VectoriSHT with their reciprocals)."""
    nside, lmax, mmax = 8, 17, 13
    o = VectorSHTOracle.healpix(nside, lmax, mmax, norm=norm)
    rng = np.random.default_rng(33)
    f = rng.standard_normal((2, 2, o.grid.npix)).astype(np.float32)
    fwd = cuhpx_amd.VectorSHT(nside, lmax=lmax, mmax=mmax, norm=norm)
    inv = cuhpx_amd.VectoriSHT(nside, lmax=lmax, mmax=mmax, norm=norm)
    c = fwd(torch.from_numpy(f).cuda()).cpu().numpy()
    assert rel_l2(c, o.forward(f.astype(np.float64))) < 1e-5
    a = vector_coeffs(lmax, mmax, 2, 34).astype(np.complex64)
    g = inv(torch.from_numpy(a).cuda()).cpu().numpy()
    assert rel_l2(g, o.inverse(a.astype(np.complex128))) < 1e-5
    # and the scaling is not the identity
    ortho = cuhpx_amd.VectorSHT(nside, lmax=lmax, mmax=mmax)(torch.from_numpy(f).cuda()).cpu().numpy()
    assert rel_l2(c, ortho) > 0.5

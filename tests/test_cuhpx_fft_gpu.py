"""The cuhpx_fft-compatible per-ring FFTs vs the oracle's per-ring DFT."""

import numpy as np
import pytest
import torch

from cuhpx_amd import cuhpx_fft
from oracle import sht as S
from oracle.geometry import RingGrid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nside,L", [(2, 5), (4, 9), (4, 30), (8, 17), (32, 65), (64, 129)])
def test_rfft_irfft_match_oracle(nside, L):
    g = RingGrid.healpix(nside, "none")
    rng = np.random.default_rng(nside)
    f = rng.standard_normal((3, g.npix)).astype(np.float32)
    ftm = cuhpx_fft.healpix_rfft_batch(torch.from_numpy(f).cuda(), L, nside)
    assert ftm.shape == (3, 4 * nside - 1, L) and ftm.dtype == torch.complex64
    ref = S.ring_rfft(f.astype(np.float64), g, L)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(ftm.cpu().numpy(), ref, atol=3e-6 * scale)
    spec = (rng.standard_normal((3, g.nring, L)) + 1j * rng.standard_normal((3, g.nring, L))).astype(np.complex64)
    spec_t = torch.from_numpy(spec).cuda()
    keep = spec_t.clone()
    back = cuhpx_fft.healpix_irfft_batch(spec_t, L, nside)
    assert torch.equal(spec_t, keep)  # input untouched
    ref = S.ring_irfft(spec.astype(np.complex128), g, L)
    np.testing.assert_allclose(back.cpu().numpy(), ref, atol=3e-6 * np.abs(ref).max())


def test_class_variants_one_map():
    nside, L = 8, 17
    f = torch.randn(12 * nside * nside, device="cuda")
    a = cuhpx_fft.healpix_rfft_class(f, L, nside)
    assert a.shape == (4 * nside - 1, L)
    torch.testing.assert_close(a, cuhpx_fft.healpix_rfft_batch(f[None], L, nside)[0])
    back = cuhpx_fft.healpix_irfft_class(a, L, nside)
    assert back.shape == f.shape

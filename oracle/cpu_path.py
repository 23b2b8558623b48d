"""Restatement of the reference's CPU torch SHT path (the timing baseline).

Test infrastructure only (see ``oracle/__init__.py``).

Follows cuhpx/hpx_sht.py:
  * ``healpix_rfft_torch``   hpx_sht.py:37-55  (per-ring torch.fft.rfft loop)
  * ``healpix_irfft_torch``  hpx_sht.py:58-76  (per-ring irfft(norm="forward"))
  * ``SHT.forward``          hpx_sht.py:214-236 (einsum with FP64 weights cast
                             to the input dtype on every call)
  * ``iSHT.forward``         hpx_sht.py:268-283
with the Legendre table of sht_tools.py:192-229 (``legpoly``, norm="ortho").
As in the reference the torch path is 1-D only (it slices dim 0), so batched
work loops over maps.
"""

import numpy as np
import torch

from .geometry import RingGrid
from .legendre import legpoly


class CPUTorchSHT:
    """SHT and iSHT of the reference CPU path, for one nside."""

    def __init__(self, nside, lmax=None, mmax=None, quad_weights="ring"):
        self.nside = nside
        self.lmax = lmax or 4 * nside - 1
        self.mmax = mmax or 2 * nside + 1
        self.grid = RingGrid.healpix(nside, quad_weights)
        pct = legpoly(self.mmax, self.lmax, np.cos(self.grid.theta))
        self.pct = torch.from_numpy(pct)                      # [mmax, lmax, nring] f64
        w = torch.from_numpy(self.grid.weight)
        self.weights = torch.einsum("mlk,k->mlk", self.pct, w)  # hpx_sht.py:211

    def rfft(self, f):
        """hpx_sht.py:37-55."""
        L = self.mmax
        ctype = torch.complex64 if f.dtype == torch.float32 else torch.complex128
        ftm = torch.zeros((self.grid.nring, L), dtype=ctype)
        for t in range(self.grid.nring):
            n = int(self.grid.nphi[t])
            o = int(self.grid.offset[t])
            k = min(n // 2 + 1, L)
            chunk = torch.fft.rfft(f[o:o + n], norm="backward")
            ftm[t, :k] = chunk[:k]
            ftm[t, :] *= torch.exp(-1j * torch.arange(L) * float(self.grid.phi0[t]))
        return ftm

    def irfft(self, ftm):
        """hpx_sht.py:58-76."""
        ftype = torch.float32 if ftm.dtype == torch.complex64 else torch.float64
        f = torch.zeros(self.grid.npix, dtype=ftype)
        L = self.mmax
        for t in range(self.grid.nring):
            ftm[t, :] *= torch.exp(1j * torch.arange(L) * float(self.grid.phi0[t]))
            n = int(self.grid.nphi[t])
            o = int(self.grid.offset[t])
            f[o:o + n] = torch.fft.irfft(ftm[t, :], n=n, norm="forward")
        return f

    def forward(self, x):
        """SHT.forward, hpx_sht.py:214-236 (1-D input)."""
        x = torch.view_as_real(self.rfft(x))
        out = torch.zeros((self.lmax, self.mmax, 2), dtype=x.dtype)
        w = self.weights.to(x.dtype)
        out[..., 0] = torch.einsum("...km,mlk->...lm", x[..., :self.mmax, 0], w)
        out[..., 1] = torch.einsum("...km,mlk->...lm", x[..., :self.mmax, 1], w)
        return torch.view_as_complex(out)

    def inverse(self, a):
        """iSHT.forward, hpx_sht.py:268-283 (2-D input)."""
        x = torch.view_as_real(a)
        p = self.pct.to(x.dtype)
        rl = torch.einsum("...lm,mlk->...km", x[..., 0], p)
        im = torch.einsum("...lm,mlk->...km", x[..., 1], p)
        return self.irfft(torch.view_as_complex(torch.stack((rl, im), -1).contiguous()))

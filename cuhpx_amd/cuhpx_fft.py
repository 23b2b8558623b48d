"""Per-ring HEALPix real FFTs with the reference's cuhpx_fft interface.

``healpix_rfft_batch(f, L, nside)`` / ``healpix_rfft_class(f, L, nside)``:
real [..., 12 nside^2] (RING) -> complex [..., 4 nside - 1, L] with
ftm[t, m] = e^{-i m phi0_t} rfft(ring t)[m] for m < min(L, N_t/2 + 1), zero above
(src/harmonic_transform/hpx_fft.cpp:21-84, 152-193).

``healpix_irfft_batch(ftm, L, nside)`` / ``healpix_irfft_class(ftm, L, nside)``:
complex [..., 4 nside - 1, L] -> real [..., 12 nside^2], per ring
irfft(ftm[t] e^{i m phi0_t}, n = N_t, norm = "forward")
(hpx_fft.cpp:86-149, 195-238).  The input is not modified (the reference's
kernels multiply the phase into it in place, hpx_fft.cpp:135, 225).

Both run the ring-pair FFT kernels of libhpxmi.so (Stockham in LDS, Bluestein
for non-power-of-two rings) on an FFT-only plan (no Legendre table).  As in the
reference (hpx_fft.cpp:45), float64 maps give complex128 spectra and complex128
spectra float64 maps, computed in FP64.
"""

import torch

from . import _lib
from .hpx_sht import ShtPlan


def _plan(L, nside, device):
    return ShtPlan.healpix(nside, 0, L, "none", device)


def healpix_rfft_batch(f: torch.Tensor, L: int, nside: int) -> torch.Tensor:
    if f.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"float32 or float64 maps only (got {f.dtype})")
    npix = 12 * nside * nside
    if f.shape[-1] != npix:
        raise RuntimeError(f"last dimension must be {npix}")
    if not f.is_cuda:
        raise RuntimeError("needs a GPU tensor")
    plan = _plan(L, nside, f.device)
    rows = f.reshape(-1, npix).contiguous()
    f64 = f.dtype == torch.float64
    if f64:
        plan.enable_f64()
    out = torch.empty((rows.shape[0], 4 * nside - 1, L), dtype=torch.complex128 if f64 else torch.complex64,
                      device=f.device)
    fn = _lib.lib().hpx_ring_rfft_f64 if f64 else _lib.lib().hpx_ring_rfft
    with torch.cuda.device(f.device):
        rc = fn(plan.handle, _lib.ptr(rows), _lib.ptr(out), rows.shape[0], _lib.stream_ptr(f.device))
    _lib.check(rc, "healpix_rfft")
    return out.reshape(*f.shape[:-1], 4 * nside - 1, L)


def healpix_irfft_batch(ftm: torch.Tensor, L: int, nside: int) -> torch.Tensor:
    if ftm.dtype not in (torch.complex64, torch.complex128):
        raise TypeError(f"complex64 or complex128 spectra only (got {ftm.dtype})")
    nring = 4 * nside - 1
    if tuple(ftm.shape[-2:]) != (nring, L):
        raise RuntimeError(f"last dimensions must be ({nring}, {L})")
    if not ftm.is_cuda:
        raise RuntimeError("needs a GPU tensor")
    plan = _plan(L, nside, ftm.device)
    rows = ftm.reshape(-1, nring, L).contiguous()
    npix = 12 * nside * nside
    f64 = ftm.dtype == torch.complex128
    if f64:
        plan.enable_f64()
    out = torch.empty((rows.shape[0], npix), dtype=torch.float64 if f64 else torch.float32, device=ftm.device)
    fn = _lib.lib().hpx_ring_irfft_f64 if f64 else _lib.lib().hpx_ring_irfft
    with torch.cuda.device(ftm.device):
        rc = fn(plan.handle, _lib.ptr(rows), _lib.ptr(out), rows.shape[0], _lib.stream_ptr(ftm.device))
    _lib.check(rc, "healpix_irfft")
    return out.reshape(*ftm.shape[:-2], npix)


def healpix_rfft_class(f: torch.Tensor, L: int, nside: int) -> torch.Tensor:
    """1-D map -> [4 nside - 1, L] (hpx_fft.cpp:152-193)."""
    return healpix_rfft_batch(f, L, nside)


def healpix_irfft_class(ftm: torch.Tensor, L: int, nside: int) -> torch.Tensor:
    """[4 nside - 1, L] -> 1-D map (hpx_fft.cpp:195-238)."""
    return healpix_irfft_batch(ftm, L, nside)

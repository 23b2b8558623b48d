"""FP64 forward / inverse SHT on an iso-latitude ring grid, and exact adjoints.

Test infrastructure only (see ``oracle/__init__.py``).

The operator is the one the reference computes (cuhpx/hpx_sht.py):

forward (SHT, hpx_sht.py:214-236 with healpix_rfft_torch 37-55):
    F_t[m] = e^{-i m phi0_t} * sum_j f[off_t + j] e^{-2 pi i m j / N_t},
             for m < K_t = min(mmax, N_t/2 + 1), and 0 for K_t <= m < mmax
    a[l, m] = sum_t w_t P_lm(theta_t) F_t[m]

inverse (iSHT, hpx_sht.py:268-283 with healpix_irfft_torch 58-76):
    G_t[m] = sum_l P_lm(theta_t) a[l, m]
    f[off_t + j] = irfft(G_t[:K_t] * e^{i m phi0_t}, n=N_t, norm="forward")[j]
                 = sum_{m < K_t} c_m Re(G_t[m] e^{i m (phi0_t + 2 pi j / N_t)}),
      c_m = 1 for m = 0 and m = N_t/2, 2 otherwise (the DC and Nyquist
      imaginary parts are dropped, as numpy/torch irfft do).

The exact adjoints (complex values taken as pairs of reals, which is
PyTorch's gradient convention) are

    d f = w_t * sum_{m<K_t} Re(H_t[m] e^{i m phi_tj}),  H_t[m] = sum_l P a_bar[l, m]
    d a = sum_t P_lm c_m sum_j g_j e^{-i m phi_tj}

The reference's hand-written backward passes (hpx_sht.py:564-583, 613-641)
differ from these by 0.5x / 2x on m = 0 and on the Nyquist mode; ``compat``
variants restate them.
"""

import numpy as np

from .geometry import RingGrid
from .legendre import legpoly, legpoly_scaled_lrows


def ring_rfft(f, grid, mmax):
    """Per-ring truncated, phase-shifted DFT: f[..., npix] -> F[..., nring, mmax]."""
    f = np.asarray(f, dtype=np.float64)
    lead = f.shape[:-1]
    out = np.zeros(lead + (grid.nring, mmax), dtype=np.complex128)
    m = np.arange(mmax)
    for t in range(grid.nring):
        n = int(grid.nphi[t])
        o = int(grid.offset[t])
        k = min(mmax, n // 2 + 1)
        spec = np.fft.rfft(f[..., o:o + n], axis=-1)
        out[..., t, :k] = spec[..., :k]
        out[..., t, :] *= np.exp(-1j * m * grid.phi0[t])
    return out


def ring_irfft(g, grid, mmax, cmode="irfft"):
    """Per-ring synthesis: G[..., nring, mmax] -> f[..., npix].

    cmode="irfft": numpy irfft(norm="forward") semantics (c_m = 1, 2, ..., 2, 1).
    cmode="ones":  sum_{m<K} Re(G e^{i m phi}) (the operator of the exact
                   SHT adjoint).
    """
    g = np.asarray(g, dtype=np.complex128)
    lead = g.shape[:-2]
    f = np.zeros(lead + (grid.npix,), dtype=np.float64)
    m = np.arange(mmax)
    for t in range(grid.nring):
        n = int(grid.nphi[t])
        o = int(grid.offset[t])
        k = min(mmax, n // 2 + 1)
        spec = g[..., t, :k] * np.exp(1j * m[:k] * grid.phi0[t])
        if cmode == "irfft":
            f[..., o:o + n] = np.fft.irfft(spec, n=n, axis=-1, norm="forward")
        elif cmode == "ones":
            c = np.full(k, 0.5)
            c[0] = 1.0
            if n % 2 == 0 and k == n // 2 + 1:
                c[n // 2] = 1.0
            f[..., o:o + n] = np.fft.irfft(spec * c, n=n, axis=-1, norm="forward")
        else:
            raise ValueError(cmode)
    return f


def cm_factors(grid, mmax):
    """c_{t,m}: 1 at m = 0 and m = N_t/2, 2 for other m < K_t, 0 beyond."""
    c = np.zeros((grid.nring, mmax))
    for t in range(grid.nring):
        n = int(grid.nphi[t])
        k = min(mmax, n // 2 + 1)
        c[t, :k] = 2.0
        c[t, 0] = 1.0
        if k == n // 2 + 1:
            c[t, n // 2] = 1.0
    return c


class SHTOracle:
    """FP64 forward / inverse SHT and adjoints on a ring grid."""

    def __init__(self, grid, lmax, mmax, pct=None, streamed=False):
        self.grid = grid
        self.lmax = lmax
        self.mmax = mmax
        self.streamed = streamed
        if pct is None and not streamed:
            pct = legpoly(mmax, lmax, np.cos(grid.theta))
        self.pct = pct  # [mmax, lmax, nring]; None when streamed

    @classmethod
    def healpix(cls, nside, lmax=None, mmax=None, quad_weights="ring", streamed=False):
        """``streamed``: generate P one m at a time (scaled recurrence) instead
        of holding the [mmax, lmax, nring] table, for large grids."""
        lmax = lmax or 4 * nside - 1
        mmax = mmax or 2 * nside + 1
        return cls(RingGrid.healpix(nside, quad_weights), lmax, mmax, streamed=streamed)

    _LBLOCK = 32

    def _lblocks(self):
        """(l0, Q) with Q[m, l - l0, k] = P_lm(x_k) for a block of l; zero for m > l."""
        nx = self.grid.nring
        buf = None
        l0 = 0
        for l, q in legpoly_scaled_lrows(self.mmax, self.lmax, np.cos(self.grid.theta)):
            if buf is None:
                buf = np.zeros((self.mmax, self._LBLOCK, nx))
                l0 = l
            buf[:q.shape[0], l - l0, :] = q
            if l - l0 + 1 == self._LBLOCK or l == self.lmax - 1:
                yield l0, buf[:, :l - l0 + 1, :]
                buf = None

    def legendre_analysis(self, ftm):
        """a[..., l, m] = sum_k P[m, l, k] ftm[..., k, m] (einsum '...km,mlk->...lm')."""
        ftm = np.asarray(ftm, dtype=np.complex128)
        out = np.zeros(ftm.shape[:-2] + (self.lmax, self.mmax), dtype=np.complex128)
        if self.streamed:  # blocks of l, one batched product over m per block
            lead = ftm.shape[:-2]
            fm = np.moveaxis(ftm.reshape((-1,) + ftm.shape[-2:]), -1, 0)  # [m, F, k]
            for l0, q in self._lblocks():
                blk = np.matmul(fm, np.swapaxes(q, 1, 2))  # [m, F, lb]
                out.reshape((-1, self.lmax, self.mmax))[:, l0:l0 + q.shape[1], :] = np.moveaxis(blk, 0, -1)
            return out.reshape(lead + (self.lmax, self.mmax))
        for m in range(self.mmax):
            out[..., :, m] = ftm[..., :, m] @ self.pct[m].T
        return out

    def legendre_synthesis(self, a):
        """g[..., k, m] = sum_l P[m, l, k] a[..., l, m] (einsum '...lm,mlk->...km')."""
        a = np.asarray(a, dtype=np.complex128)
        out = np.zeros(a.shape[:-2] + (self.grid.nring, self.mmax), dtype=np.complex128)
        if self.streamed:
            lead = a.shape[:-2]
            am = np.moveaxis(a.reshape((-1,) + a.shape[-2:]), -1, 0)  # [m, F, l]
            acc = np.zeros((self.mmax, am.shape[1], self.grid.nring), dtype=np.complex128)
            for l0, q in self._lblocks():
                acc += np.matmul(am[:, :, l0:l0 + q.shape[1]], q.astype(np.complex128))
            return np.moveaxis(acc, 0, -1).reshape(lead + (self.grid.nring, self.mmax))
        for m in range(self.mmax):
            out[..., :, m] = a[..., :, m] @ self.pct[m]
        return out

    def forward_and_inverse(self, f, a):
        """(forward(f), inverse(a)) with one sweep over the table: a streamed
        oracle generates P once for both directions instead of twice."""
        if not self.streamed:
            return self.forward(f), self.inverse(a)
        ftm = ring_rfft(f, self.grid, self.mmax) * self.grid.weight[:, None]
        a = np.asarray(a, dtype=np.complex128)
        lead_f, lead_a = ftm.shape[:-2], a.shape[:-2]
        fm = np.moveaxis(ftm.reshape((-1,) + ftm.shape[-2:]), -1, 0)  # [m, F, k]
        am = np.moveaxis(a.reshape((-1,) + a.shape[-2:]), -1, 0)      # [m, F, l]
        alm = np.zeros((fm.shape[1], self.lmax, self.mmax), dtype=np.complex128)
        acc = np.zeros((self.mmax, am.shape[1], self.grid.nring), dtype=np.complex128)
        for l0, q in self._lblocks():
            lb = q.shape[1]
            alm[:, l0:l0 + lb, :] = np.moveaxis(np.matmul(fm, np.swapaxes(q, 1, 2)), 0, -1)
            acc += np.matmul(am[:, :, l0:l0 + lb], q.astype(np.complex128))
        g = np.moveaxis(acc, 0, -1).reshape(lead_a + (self.grid.nring, self.mmax))
        return alm.reshape(lead_f + (self.lmax, self.mmax)), ring_irfft(g, self.grid, self.mmax, "irfft")

    def forward(self, f):
        """a[..., lmax, mmax] = sum_t w_t P F_t (hpx_sht.py:214-236)."""
        ftm = ring_rfft(f, self.grid, self.mmax)
        return self.legendre_analysis(ftm * self.grid.weight[:, None])

    def inverse(self, a):
        """f = irfft(sum_l P a) per ring (hpx_sht.py:268-283)."""
        return ring_irfft(self.legendre_synthesis(a), self.grid, self.mmax, "irfft")

    def forward_adjoint(self, ga):
        """Exact gradient of <forward(f), ga> w.r.t. f (real pairing)."""
        h = self.legendre_synthesis(ga)
        f = ring_irfft(h, self.grid, self.mmax, "ones")
        return f * self._pixel_weight()

    def inverse_adjoint(self, gf):
        """Exact gradient of <inverse(a), gf> w.r.t. a (real pairing)."""
        ftm = ring_rfft(gf, self.grid, self.mmax) * cm_factors(self.grid, self.mmax)
        return self.legendre_analysis(ftm)

    def forward_adjoint_compat(self, ga):
        """The reference's SHTFunction.backward: W * irfft(P^T ga), W = w/2
        (hpx_sht.py:564-583, sht_tools.py:307-316)."""
        h = self.legendre_synthesis(ga)
        f = ring_irfft(h, self.grid, self.mmax, "irfft")
        return f * self._pixel_weight() / 2

    def inverse_adjoint_compat(self, gf):
        """The reference's iSHTFunction.backward: (P w) rfft(gf / W)
        (hpx_sht.py:613-641)."""
        wpix = self._pixel_weight() / 2
        ftm = ring_rfft(np.asarray(gf) / wpix, self.grid, self.mmax)
        return self.legendre_analysis(ftm * self.grid.weight[:, None])

    def _pixel_weight(self):
        return np.repeat(self.grid.weight, self.grid.nphi)


def dft_forward_bruteforce(f, grid, mmax):
    """Direct O(N * mmax) evaluation of F_t[m] (no FFT), for tiny grids."""
    f = np.asarray(f, dtype=np.float64)
    out = np.zeros(f.shape[:-1] + (grid.nring, mmax), dtype=np.complex128)
    for t in range(grid.nring):
        n = int(grid.nphi[t])
        o = int(grid.offset[t])
        k = min(mmax, n // 2 + 1)
        phi = grid.phi0[t] + 2 * np.pi * np.arange(n) / n
        for m in range(k):
            out[..., t, m] = (f[..., o:o + n] * np.exp(-1j * m * phi)).sum(-1)
    return out


def dft_inverse_bruteforce(g, grid, mmax):
    """Direct sum_{m<K} c_m Re(G e^{i m phi}), for tiny grids."""
    g = np.asarray(g, dtype=np.complex128)
    f = np.zeros(g.shape[:-2] + (grid.npix,), dtype=np.float64)
    c = cm_factors(grid, mmax)
    for t in range(grid.nring):
        n = int(grid.nphi[t])
        o = int(grid.offset[t])
        phi = grid.phi0[t] + 2 * np.pi * np.arange(n) / n
        for m in range(mmax):
            if c[t, m] == 0:
                continue
            f[..., o:o + n] += c[t, m] * np.real(g[..., t, m, None] * np.exp(1j * m * phi))
    return f

"""FP64 vector (spheroidal / toroidal) SHT on a ring grid.

Test infrastructure only (see ``oracle/__init__.py``).

Restates VectorSHT / VectoriSHT (cuhpx/hpx_sht.py:286-444) and the table
they precompute, _precompute_dlegpoly (cuhpx/sht_tools.py:267-304):

    dpct[0, m, l] = d/dtheta P^m_l          (Condon-Shortley sign on odd m)
    dpct[1, m, l] = m P^m_l / sin(theta)     (1 <= m <= l; zero at m = 0)

built from the CS-free orthonormal P of ``legendre.legpoly`` by the
ladder identities the reference uses.  With ``csphase=True`` (the
reference's default) the two rows are the theta-derivative and the
m / sin(theta) multiple of P_CS = (-1)^m P; tests pin both identities.

forward, v[..., 2, npix] -> c[..., 2, lmax, mmax] (hpx_sht.py:333-374):
    F_j = ring_rfft(v_j)                      (no quadrature weight here)
    W0 = dpct[0] w_k / (l (l + 1)),  W1 = -dpct[1] w_k / (l (l + 1))   (l = 0: 1)
    S = W0 F0 + i W1 F1,   T = i W1 F0 - W0 F1   (products '...km,mlk->...lm')

inverse, c[..., 2, lmax, mmax] -> v[..., 2, npix] (hpx_sht.py:408-444):
    s = D0 a0 + i D1 a1,   t = i D1 a0 - D0 a1   (D = dpct, '...lm,mlk->...km')
    v = ring_irfft((s, t))

So the inverse of (a, 0) is the gradient (d_theta g, d_phi g / sin theta) of
the scalar synthesis g of a with CS-phase harmonics, and the inverse of (0, b)
is (d_phi h / sin theta, -d_theta h).
"""

import numpy as np

from .geometry import RingGrid
from .legendre import legpoly
from .sht import ring_irfft, ring_rfft


def norm_scale(lmax, norm="ortho", inverse=False):
    """Per-degree factor legpoly(norm=..., inverse=...) applies to its rows
    (sht_tools.py:196-221): 1 for "ortho"; sqrt(4 pi) for any other norm,
    times (2L+1)^(-1/2) for "schmidt"; reciprocals when ``inverse``."""
    L = np.arange(lmax, dtype=np.float64)
    f = np.ones(lmax)
    if norm != "ortho":
        f *= np.sqrt(4 * np.pi)
        if norm == "schmidt":
            f /= np.sqrt(2 * L + 1)
    return 1.0 / f if inverse else f


def dlegpoly(mmax, lmax, x, csphase=True, norm="ortho", inverse=False):
    """dpct[2, mmax, lmax, len(x)] of sht_tools.py:267-304, vectorised over m."""
    x = np.asarray(x, dtype=np.float64)
    pct = legpoly(mmax + 1, lmax + 1, x)  # P^m_l for m <= mmax, l <= lmax
    pct = pct * norm_scale(lmax + 1, norm, inverse)[None, :, None]
    d = np.zeros((2, mmax, lmax, len(x)), dtype=np.float64)
    for l in range(lmax):
        # m = 0 (sht_tools.py:279)
        d[0, 0, l] = -np.sqrt(l * (l + 1)) * pct[1, l]
        # 0 < m < l (282-285)
        m = np.arange(1, min(l, mmax))
        if len(m):
            d[0, m, l] = 0.5 * (np.sqrt((l + m) * (l - m + 1))[:, None] * pct[m - 1, l]
                                - np.sqrt((l - m) * (l + m + 1))[:, None] * pct[m + 1, l])
        # m == l (288-289)
        if mmax > l and l > 0:
            d[0, l, l] = np.sqrt(l / 2) * pct[l - 1, l]
        # 1 <= m <= l: m P / sin(theta) by the ladder identity (292-301)
        m = np.arange(1, min(l + 1, mmax))
        if len(m):
            d[1, m, l] = 0.5 * np.sqrt((2 * l + 1) / (2 * l + 3)) * (
                np.sqrt((l - m + 1) * (l - m + 2))[:, None] * pct[m - 1, l + 1]
                + np.sqrt((l + m + 1) * (l + m + 2))[:, None] * pct[m + 1, l + 1])
    if csphase:  # 302-304
        d[:, 1::2] *= -1
    return d


def _contract_lm(F, W):
    """out[..., l, m] = sum_k W[m, l, k] F[..., k, m]  ('...km,mlk->...lm')."""
    out = np.zeros(F.shape[:-2] + (W.shape[1], W.shape[0]), dtype=np.complex128)
    for m in range(W.shape[0]):
        out[..., :, m] = F[..., :, m] @ W[m].T
    return out


def _contract_km(a, W):
    """out[..., k, m] = sum_l W[m, l, k] a[..., l, m]  ('...lm,mlk->...km')."""
    out = np.zeros(a.shape[:-2] + (W.shape[2], W.shape[0]), dtype=np.complex128)
    for m in range(W.shape[0]):
        out[..., :, m] = a[..., :, m] @ W[m]
    return out


class VectorSHTOracle:
    """FP64 VectorSHT / VectoriSHT on a ring grid."""

    def __init__(self, grid, lmax, mmax, csphase=True, norm="ortho"):
        self.grid = grid
        self.lmax = lmax
        self.mmax = mmax
        x = np.cos(grid.theta)
        self.dpct = dlegpoly(mmax, lmax, x, csphase, norm)  # VectorSHT's table
        # VectoriSHT's (hpx_sht.py:396: inverse=True); the same for "ortho"
        self.dpct_inv = self.dpct if norm == "ortho" else dlegpoly(mmax, lmax, x, csphase, norm, inverse=True)

    @classmethod
    def healpix(cls, nside, lmax=None, mmax=None, quad_weights="ring", csphase=True, norm="ortho"):
        lmax = lmax or 4 * nside - 1
        mmax = mmax or 2 * nside + 1
        return cls(RingGrid.healpix(nside, quad_weights), lmax, mmax, csphase, norm)

    def norm_factor(self):
        """1 / (l (l + 1)) with 1 at l = 0 (hpx_sht.py:324-326)."""
        l = np.arange(self.lmax, dtype=np.float64)
        nf = np.ones(self.lmax)
        nf[1:] = 1.0 / (l[1:] * (l[1:] + 1))
        return nf

    def forward(self, v):
        """v[..., 2, npix] -> (S, T) coefficients [..., 2, lmax, mmax]."""
        F = ring_rfft(np.asarray(v, dtype=np.float64), self.grid, self.mmax)
        scale = self.grid.weight[None, None, :] * self.norm_factor()[None, :, None]
        W0 = self.dpct[0] * scale
        W1 = -self.dpct[1] * scale
        F0, F1 = F[..., 0, :, :], F[..., 1, :, :]
        S = _contract_lm(F0, W0) + 1j * _contract_lm(F1, W1)
        T = 1j * _contract_lm(F0, W1) - _contract_lm(F1, W0)
        return np.stack([S, T], axis=-3)

    def inverse(self, c):
        """(S, T) coefficients [..., 2, lmax, mmax] -> v[..., 2, npix]."""
        c = np.asarray(c, dtype=np.complex128)
        a0, a1 = c[..., 0, :, :], c[..., 1, :, :]
        D0, D1 = self.dpct_inv[0], self.dpct_inv[1]
        s = _contract_km(a0, D0) + 1j * _contract_km(a1, D1)
        t = 1j * _contract_km(a0, D1) - _contract_km(a1, D0)
        return ring_irfft(np.stack([s, t], axis=-3), self.grid, self.mmax, "irfft")


def scalar_field_at(a, theta, phi, csphase=True):
    """g(theta, phi) = sum_m c_m Re(sum_l P^m_l(cos theta) a[l, m] e^{i m phi})
    with c_0 = 1, c_m = 2 (the real field whose ring spectra a synthesises),
    P with the CS sign on odd m when ``csphase``; brute force at arbitrary
    points, for tiny band limits."""
    a = np.asarray(a, dtype=np.complex128)
    lmax, mmax = a.shape
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    P = legpoly(mmax, lmax, np.cos(theta))  # [m, l, n]
    if csphase:
        P = P * ((-1.0) ** np.arange(mmax))[:, None, None]
    G = np.einsum("mln,lm->mn", P, a)
    c = np.full(mmax, 2.0)
    c[0] = 1.0
    return np.einsum("m,mn->n", c, np.real(G * np.exp(1j * np.arange(mmax)[:, None] * phi[None, :])))

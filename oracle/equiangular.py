"""Lat/lon grids for the regridding leg (FP64 oracle).

Test infrastructure only (see ``oracle/__init__.py``).

The reference's lat/lon leg is torch_harmonics.RealSHT / InverseRealSHT
(cuhpx/hpx_regrid.py:68-71, 84-86; torch_harmonics is not vendored and not
installed here).  This module states the operator from its published
definition: nodes theta_j = pi j / (nlat - 1) with Clenshaw-Curtis weights for
'equiangular', Gauss-Legendre nodes for 'legendre-gauss', Gauss-Lobatto-
Legendre nodes (+-1 and the roots of P'_{nlat-1}) for 'lobatto'; the forward
transform a_lm = sum_j w_j (2 pi / nlon) P_lm(theta_j) sum_k f_jk e^{-i m phi_k}
with the Condon-Shortley phase (-1)^m when csphase is set, and the inverse
irfft(sum_l P a, n=nlon, norm="forward").

The Clenshaw-Curtis weights here are obtained by solving the Chebyshev moment
equations sum_j w_j cos(k theta_j) = int_{-1}^{1} T_k(x) dx (k < nlat), which
is independent of the closed-form cosine sum the product uses; the Lobatto
weights likewise solve sum_j w_j P_k(x_j) = 2 [k = 0] (k < nlat) instead of
using the closed form 2 / (n (n-1) P_{n-1}(x)^2).  All three rules are pinned
by textbook weights at small n and by exactness on polynomials
(tests/test_oracle_latlon.py).

parity unpinned: torch_harmonics' own numerics (only the round trip and the
quadrature's exactness are pinned).
"""

import numpy as np

from .geometry import RingGrid


def clenshaw_curtis_by_moments(n):
    theta = np.pi * np.arange(n) / (n - 1)
    k = np.arange(n)
    T = np.cos(k[:, None] * theta[None, :])
    with np.errstate(divide="ignore"):
        mom = np.where(k % 2 == 0, 2.0 / (1.0 - k.astype(np.float64) ** 2), 0.0)
    w = np.linalg.solve(T, mom)
    return np.cos(theta), w


def gauss_lobatto_by_moments(n):
    leg = np.polynomial.legendre.Legendre
    inner = leg.basis(n - 1).deriv().roots().real if n > 2 else np.array([])
    x = np.concatenate([[1.0], np.sort(inner)[::-1], [-1.0]])
    V = np.array([leg.basis(k)(x) for k in range(n)])
    rhs = np.zeros(n)
    rhs[0] = 2.0
    return x, np.linalg.solve(V, rhs)


def latlon_grid(nlat, nlon, grid="equiangular", csphase=True):
    if grid == "equiangular":
        x, w = clenshaw_curtis_by_moments(nlat)
    elif grid == "legendre-gauss":
        x, w = np.polynomial.legendre.leggauss(nlat)
        x, w = x[::-1], w[::-1]
    elif grid == "lobatto":
        x, w = gauss_lobatto_by_moments(nlat)
    else:
        raise ValueError(grid)
    theta = np.arccos(np.clip(x, -1, 1))
    phi0 = np.full(nlat, np.pi if csphase else 0.0)
    return RingGrid(theta, np.full(nlat, nlon), phi0, w * 2 * np.pi / nlon)

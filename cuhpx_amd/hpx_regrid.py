"""Regridding between HEALPix and lat/lon grids through the SHT.

Drop-in for cuhpx/hpx_regrid.py:21-91: ``Grid(grid, grid_size, norm, csphase)``
and ``Regridding(src_grid, dest_grid, lmax=None, mmax=None, device=None,
enable_cuda=True).execute(f) = isht(sht(f))`` with the reference's default
band limits (lmax = min over grids of 2*nside+1 or nlat; mmax = min of
2*nside+1 or nlon//2+1, hpx_regrid.py:50-58).

The HEALPix legs are SHTCUDA / iSHTCUDA.  The reference delegates the lat/lon
legs to torch_harmonics.RealSHT / InverseRealSHT (hpx_regrid.py:68-71, 84-86);
here they are the same ring-grid kernels of libhpxmi.so on a lat/lon plan:
  * nodes and quadrature: 'equiangular' = Clenshaw-Curtis nodes
    theta_j = pi j / (nlat - 1) (poles included) and weights; 'legendre-gauss'
    = Gauss-Legendre; 'lobatto' = Gauss-Lobatto-Legendre.  The reference
    passes no grid type to torch_harmonics, so it always uses its default
    (equiangular) nodes; this module honours the requested grid;
  * forward: a_lm = sum_j w_j (2 pi / nlon) P_lm(theta_j) sum_k f_jk e^{-i m phi_k}
    (torch_harmonics' 2*pi*rfft(norm="forward") convention);
  * inverse: irfft(sum_l P_lm a_lm, n=nlon, norm="forward") per latitude;
  * csphase=True applies the Condon-Shortley phase (-1)^m, implemented as a
    longitude origin of pi (e^{i m pi} = (-1)^m) in the ring plan.
The HEALPix side has no Condon-Shortley phase (its csphase argument is ignored,
as in the reference), so a hpx -> lat/lon map can differ from a phase-free one
by a rotation of pi in longitude for odd m; the round trip cancels it.

enable_cuda is accepted for signature compatibility; the GPU path always runs.
"""

import numpy as np
import torch
import torch.nn as nn

from .hpx_sht import SHTCUDA, ShtPlan, SHTFunction, iSHTCUDA, iSHTFunction, _check_opts


def clenshaw_curtis(n):
    """Nodes cos(theta_j), theta_j = pi j/(n-1) (north to south), and weights
    w_j = (c_j/(n-1)) [1 - sum_{k=1}^{(n-1)//2} b_k cos(2 k theta_j)/(4k^2 - 1)],
    c_j = 1 at the poles and 2 inside, b_k = 1 for k = (n-1)/2 and 2 otherwise."""
    if n < 2:
        raise ValueError("Clenshaw-Curtis needs at least 2 nodes")
    n1 = n - 1
    theta = np.pi * np.arange(n) / n1
    k = np.arange(1, n1 // 2 + 1)
    b = np.where(2 * k == n1, 1.0, 2.0)
    s = (b[None, :] * np.cos(2 * k[None, :] * theta[:, None]) / (4 * k[None, :] ** 2 - 1)).sum(1)
    c = np.full(n, 2.0)
    c[0] = c[-1] = 1.0
    return np.cos(theta), c / n1 * (1 - s)


def gauss_legendre(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x[::-1].copy(), w[::-1].copy()


def gauss_lobatto(n):
    """Gauss-Lobatto-Legendre nodes (+-1 and the roots of P'_{n-1}), north to
    south, with weights 2 / (n (n-1) P_{n-1}(x)^2)."""
    if n < 2:
        raise ValueError("Lobatto needs at least 2 nodes")
    leg = np.polynomial.legendre
    c = np.zeros(n)
    c[-1] = 1.0  # P_{n-1}
    inner = leg.legroots(leg.legder(c)) if n > 2 else np.array([])
    dc = leg.legder(c)
    d2c = leg.legder(dc)
    for _ in range(3):  # Newton polish of the interior roots
        inner = inner - leg.legval(inner, dc) / leg.legval(inner, d2c)
    x = np.concatenate([[1.0], np.sort(inner)[::-1], [-1.0]])
    w = 2.0 / (n * (n - 1) * leg.legval(x, c) ** 2)
    return x, w


def latlon_nodes(grid, nlat):
    if grid == "equiangular":
        return clenshaw_curtis(nlat)
    if grid == "legendre-gauss":
        return gauss_legendre(nlat)
    if grid == "lobatto":
        return gauss_lobatto(nlat)
    raise ValueError("Unknown quadrature mode")


def latlon_plan(grid, nlat, nlon, lmax, mmax, csphase, device=None):
    x, w = latlon_nodes(grid, nlat)
    theta = np.arccos(np.clip(x, -1.0, 1.0))
    nphi = np.full(nlat, nlon, dtype=np.int32)
    phi0 = np.full(nlat, np.pi if csphase else 0.0)
    weights = w * (2.0 * np.pi / nlon)
    return ShtPlan.rings(nphi, theta, phi0, weights, lmax, mmax, device=device,
                         key=("latlon", grid, nlat, nlon, bool(csphase)))


class _LatLonBase(nn.Module):
    def __init__(self, nlat, nlon, lmax=None, mmax=None, grid="equiangular", norm="ortho", csphase=True,
                 precision="fp32", adjoint="exact"):
        super().__init__()
        if norm != "ortho":
            raise NotImplementedError("only norm='ortho' is implemented")
        _check_opts(precision, adjoint)
        if not torch.cuda.is_available():
            raise RuntimeError("CUDA device is not available. This class requires a GPU.")
        self.nlat, self.nlon = nlat, nlon
        self.grid, self.norm, self.csphase = grid, norm, csphase
        self.lmax = lmax or nlat
        self.mmax = mmax or (nlon // 2 + 1)
        self.precision, self.adjoint = precision, adjoint
        self._plan = latlon_plan(grid, nlat, nlon, self.lmax, self.mmax, csphase)

    def plan_for(self, device):
        if device.type != "cuda":
            raise RuntimeError("lat/lon SHT needs a GPU tensor")
        if self._plan.device.index != device.index:
            self._plan = latlon_plan(self.grid, self.nlat, self.nlon, self.lmax, self.mmax, self.csphase,
                                     device)
        return self._plan


class LatLonSHT(_LatLonBase):
    """Forward SHT of real [..., nlat, nlon] maps -> complex [..., lmax, mmax]
    (torch_harmonics.RealSHT semantics with norm='ortho')."""

    def forward(self, x):
        if torch.is_complex(x) or x.dtype not in (torch.float32, torch.float64):
            raise TypeError("LatLonSHT takes float32 or float64 maps")
        plan = self.plan_for(x.device)
        lead = x.shape[:-2]
        flat = x.reshape(*lead, self.nlat * self.nlon)
        return SHTFunction.apply(flat, plan, self.precision, self.adjoint)


class InverseLatLonSHT(_LatLonBase):
    """Inverse SHT complex [..., lmax, mmax] -> real [..., nlat, nlon]
    (torch_harmonics.InverseRealSHT semantics with norm='ortho')."""

    def forward(self, a):
        if a.dtype not in (torch.complex64, torch.complex128):
            raise TypeError("InverseLatLonSHT takes complex64 or complex128 coefficients")
        plan = self.plan_for(a.device)
        out = iSHTFunction.apply(a, plan, self.precision, self.adjoint)
        return out.reshape(*a.shape[:-2], self.nlat, self.nlon)


class Grid:
    """hpx_regrid.py:21-37."""

    def __init__(self, grid, grid_size, norm="ortho", csphase=True):
        self.grid = grid
        self.nside = None
        self.nlat = None
        self.nlon = None
        self.norm = norm
        self.csphase = csphase
        if self.grid == "healpix":
            self.nside = grid_size
        elif self.grid in ["legendre-gauss", "lobatto", "equiangular"]:
            self.nlat, self.nlon = grid_size
        else:
            raise ValueError("Unknown quadrature mode")


class Regridding:
    """hpx_regrid.py:40-91: execute(f) = isht(sht(f))."""

    def __init__(self, src_grid, dest_grid, lmax=None, mmax=None, device=None, enable_cuda=True,
                 precision="fp32", adjoint="exact"):
        self.lmax = lmax or self._determine_lmax(src_grid, dest_grid)
        self.mmax = mmax or self._determine_mmax(src_grid, dest_grid)
        self.device = device
        self.enable_cuda = enable_cuda
        self.precision = precision
        self.adjoint = adjoint
        self.sht = self._initialize_sht(src_grid, self.lmax, self.mmax).to(device)
        self.isht = self._initialize_isht(dest_grid, self.lmax, self.mmax).to(device)

    def _determine_lmax(self, src_grid, dest_grid):
        src_lmax = 2 * src_grid.nside + 1 if src_grid.nside else src_grid.nlat
        dest_lmax = 2 * dest_grid.nside + 1 if dest_grid.nside else dest_grid.nlat
        return min(src_lmax, dest_lmax)

    def _determine_mmax(self, src_grid, dest_grid):
        src_mmax = 2 * src_grid.nside + 1 if src_grid.nside else src_grid.nlon // 2 + 1
        dest_mmax = 2 * dest_grid.nside + 1 if dest_grid.nside else dest_grid.nlon // 2 + 1
        return min(src_mmax, dest_mmax)

    def _initialize_sht(self, grid, lmax, mmax):
        if grid.grid == "healpix":
            return SHTCUDA(grid.nside, lmax=lmax, mmax=mmax, norm=grid.norm, csphase=grid.csphase,
                           precision=self.precision, adjoint=self.adjoint)
        if grid.grid in ["legendre-gauss", "lobatto", "equiangular"]:
            return LatLonSHT(grid.nlat, grid.nlon, lmax=lmax, mmax=mmax, grid=grid.grid, norm=grid.norm,
                             csphase=grid.csphase, precision=self.precision, adjoint=self.adjoint)
        raise ValueError("Unknown quadrature mode")

    def _initialize_isht(self, grid, lmax, mmax):
        if grid.grid == "healpix":
            return iSHTCUDA(grid.nside, lmax=lmax, mmax=mmax, norm=grid.norm, csphase=grid.csphase,
                            precision=self.precision, adjoint=self.adjoint)
        if grid.grid in ["legendre-gauss", "lobatto", "equiangular"]:
            return InverseLatLonSHT(grid.nlat, grid.nlon, lmax=lmax, mmax=mmax, grid=grid.grid,
                                    norm=grid.norm, csphase=grid.csphase, precision=self.precision,
                                    adjoint=self.adjoint)
        raise ValueError("Unknown quadrature mode")

    def execute(self, f):
        return self.isht(self.sht(f))

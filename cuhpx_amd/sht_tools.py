"""HEALPix ring geometry and ring quadrature weights (host side, FP64).

Same names and results as cuhpx/sht_tools.py for the functions the SHT path
uses: ``read_ring_weight`` (34-44), ``apply_ring_weight`` (47-59),
``healpix_weights`` (114-140), ``p2phi_ring`` (161-176), ``nphi_ring``
(179-189), ``W_helper`` (307-316) and ``cumulative_nphi_ring`` (319-332).
The ring weights are read from ``data/ring_weights.npz``, converted once from
the HEALPix FITS tables the reference ships (see data/convert_ring_weights.py),
so astropy is not needed.  The Legendre table is not built here: the device
generates it inside the SHT plan.
"""

import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "ring_weights.npz")
_cache = {}


def read_ring_weight(nside):
    """D_weight of the 2*nside north rings."""
    if nside not in _cache:
        key = f"n{nside:05d}"
        with np.load(_DATA) as data:
            if key not in data.files:
                raise ValueError(f"no HEALPix ring weights for nside={nside} (2..8192, powers of two)")
            _cache[nside] = np.array(data[key], dtype=np.float64)
    return _cache[nside].copy()


def apply_ring_weight(nside):
    """w_t = 4 pi / npix * (1 + D[northring(t)]) for the 4*nside-1 rings."""
    w = read_ring_weight(nside)
    npix = 12 * nside * nside
    ring = np.arange(1, 4 * nside)
    northring = np.where(ring > 2 * nside, 4 * nside - ring, ring)
    return 4.0 * np.pi / npix * (1.0 + w[northring - 1])


def healpix_weights(nside, weight="ring"):
    """(z, w): z of the rings (flipped, south to north, as the reference
    returns it) and the ring quadrature weights ('ring' or uniform)."""
    t = np.arange(4 * nside - 1)
    z = np.zeros(t.shape, dtype=np.float64)
    npix = 12 * nside * nside
    if weight == "ring":
        w = apply_ring_weight(nside)
    else:
        w = 4.0 * np.pi / npix * np.ones(t.shape, dtype=np.float64)
    m1 = t < nside - 1
    m2 = (t >= nside - 1) & (t <= 3 * nside - 1)
    m3 = (t > 3 * nside - 1) & (t <= 4 * nside - 2)
    z[m1] = 1 - ((t[m1] + 1) ** 2) / (3 * nside ** 2)
    z[m2] = 4 / 3 - 2 * (t[m2] + 1) / (3 * nside)
    z[m3] = ((4 * nside - 1 - t[m3]) ** 2) / (3 * nside ** 2) - 1
    return np.flip(z), w


def nphi_ring(t, nside):
    if 0 <= t < nside - 1:
        return 4 * (t + 1)
    if nside - 1 <= t <= 3 * nside - 1:
        return 4 * nside
    if 3 * nside - 1 < t <= 4 * nside - 2:
        return 4 * (4 * nside - t - 1)
    raise ValueError(f"Ring t={t} not contained by nside={nside}")


def cumulative_nphi_ring(t, nside):
    if 0 <= t < nside:
        return 2 * t * (t + 1)
    if t < 3 * nside:
        return 2 * nside * (nside + 1) + (t - nside) * 4 * nside
    if t < 4 * nside:
        r = 4 * nside - t - 1
        return 12 * nside * nside - 2 * r * (r + 1)
    return -1


def p2phi_ring(t, p, nside):
    shift = 0.5
    if nside <= t + 1 <= 3 * nside:
        shift *= (t - nside + 2) % 2
        return np.pi / (2 * nside) * (p + shift)
    if t + 1 > 3 * nside:
        factor = np.pi / (2 * (4 * nside - t - 1))
    else:
        factor = np.pi / (2 * (t + 1))
    return factor * (p + shift)


def W_helper(w, nside):
    """Per-pixel w_t / 2 (the reference's adjoint scaling), float32."""
    nphi = np.array([nphi_ring(t, nside) for t in range(4 * nside - 1)])
    return np.repeat(np.asarray(w, dtype=np.float64) / 2, nphi).astype(np.float32)

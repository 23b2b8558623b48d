"""HEALPix ring geometry and ring quadrature weights (FP64 oracle).

Test infrastructure only (see ``oracle/__init__.py``).

Follows cuhpx/sht_tools.py:
  * ``nphi_ring``              sht_tools.py:179-189
  * ``cumulative_nphi_ring``   sht_tools.py:319-332
  * ``p2phi_ring``             sht_tools.py:161-176
  * ``healpix_weights``        sht_tools.py:114-140 (z of each ring, ring weights)
  * ``apply_ring_weight``      sht_tools.py:47-59
"""

import os

import numpy as np

_WEIGHTS_NPZ = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "cuhpx_amd", "data", "ring_weights.npz")


def nphi_ring(t, nside):
    """Pixels on ring t (0-based from the north pole); sht_tools.py:179-189."""
    if 0 <= t < nside - 1:
        return 4 * (t + 1)
    if nside - 1 <= t <= 3 * nside - 1:
        return 4 * nside
    if 3 * nside - 1 < t <= 4 * nside - 2:
        return 4 * (4 * nside - t - 1)
    raise ValueError(f"Ring t={t} not contained by nside={nside}")


def cumulative_nphi_ring(t, nside):
    """RING index of the first pixel of ring t; sht_tools.py:319-332."""
    if 0 <= t < nside:
        return 2 * t * (t + 1)
    if t < 3 * nside:
        return 2 * nside * (nside + 1) + (t - nside) * 4 * nside
    if t < 4 * nside:
        r = 4 * nside - t - 1
        return 12 * nside * nside - 2 * r * (r + 1)
    return -1


def p2phi_ring(t, p, nside):
    """Longitude of pixel p on ring t; sht_tools.py:161-176."""
    shift = 0.5
    if nside <= t + 1 <= 3 * nside:
        shift *= (t - nside + 2) % 2
        return np.pi / (2 * nside) * (p + shift)
    if t + 1 > 3 * nside:
        factor = np.pi / (2 * (4 * nside - t - 1))
    else:
        factor = np.pi / (2 * (t + 1))
    return factor * (p + shift)


def read_ring_weight(nside):
    """D_weight of the 2*nside north rings (column 0 of the FITS table)."""
    with np.load(_WEIGHTS_NPZ) as data:
        key = f"n{nside:05d}"
        if key not in data.files:
            raise ValueError(f"no ring weights for nside={nside}")
        return np.array(data[key], dtype=np.float64)


def apply_ring_weight(nside):
    """w_t = 4*pi/npix * (1 + D[northring(t)]); sht_tools.py:47-59."""
    w = read_ring_weight(nside)
    npix = 12 * nside * nside
    nrings = 4 * nside - 1
    ring = np.arange(1, nrings + 1)
    northring = np.where(ring > 2 * nside, 4 * nside - ring, ring)
    return 4.0 * np.pi / npix * (1.0 + w[northring - 1])


def healpix_weights(nside, weight="ring"):
    """(z, w) of the rings; sht_tools.py:114-140.

    As in the reference, ``z`` is returned flipped (south to north) and the
    callers take ``theta = flip(arccos(z))``, which is north to south.
    """
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


class RingGrid:
    """Iso-latitude ring layout: per ring its z, length, offset, phi0, weight.

    ``theta`` runs north to south.  For HEALPix this is what the reference's
    SHT / iSHT modules build in their constructors (hpx_sht.py:198-211 and
    645-683).
    """

    def __init__(self, theta, nphi, phi0, weight):
        self.theta = np.asarray(theta, dtype=np.float64)
        self.nphi = np.asarray(nphi, dtype=np.int64)
        self.phi0 = np.asarray(phi0, dtype=np.float64)
        self.weight = np.asarray(weight, dtype=np.float64)
        self.offset = np.concatenate([[0], np.cumsum(self.nphi)[:-1]]).astype(np.int64)
        self.npix = int(self.nphi.sum())
        self.nring = len(self.nphi)

    @classmethod
    def healpix(cls, nside, quad_weights="ring"):
        cost, w = healpix_weights(nside, quad_weights)
        theta = np.flip(np.arccos(cost))
        nring = 4 * nside - 1
        nphi = [nphi_ring(t, nside) for t in range(nring)]
        phi0 = [p2phi_ring(t, 0, nside) for t in range(nring)]
        return cls(theta, nphi, phi0, w)


def healpix_ring_of_pixel(nside):
    """Ring index of every RING-ordered pixel (brute force from offsets)."""
    nring = 4 * nside - 1
    out = np.empty(12 * nside * nside, dtype=np.int64)
    for t in range(nring):
        o = cumulative_nphi_ring(t, nside)
        out[o:o + nphi_ring(t, nside)] = t
    return out

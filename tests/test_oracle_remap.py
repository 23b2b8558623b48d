"""Pins the remap oracle (oracle/remap.py) without a GPU.

Golden values: the examples published in healpy's ``nest2ring`` /
``ring2nest`` docstrings (HEALPix's own pixel numbering); the reference's
tests compare against healpy.pixelfunc.reorder (tests/test_data_remapping.py:43-66),
which is not installed here, so these published values stand in for it.
Invariants: bijection, inverse pairs, identities of the flat layouts.
"""

import numpy as np
import pytest

from oracle import remap as R

ORIGINS = ["S", "E", "N", "W"]


def test_healpy_docstring_values():
    assert R.sigma("nest2ring", 16)[1130] == 1504
    assert R.sigma("ring2nest", 16)[1504] == 1130
    np.testing.assert_array_equal(R.sigma("nest2ring", 2)[:10], [13, 5, 4, 0, 15, 7, 6, 1, 17, 9])
    np.testing.assert_array_equal(R.sigma("ring2nest", 2)[:10], [3, 7, 11, 15, 2, 1, 6, 5, 10, 9])
    assert [int(R.sigma("nest2ring", n)[11]) for n in (1, 2, 4, 8)] == [11, 2, 12, 211]
    assert [int(R.sigma("ring2nest", n)[11]) for n in (1, 2, 4, 8)] == [11, 13, 61, 253]


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 32])
def test_ring_nest_bijection_and_inverse(nside):
    n = 12 * nside * nside
    a = R.sigma("ring2nest", nside)
    b = R.sigma("nest2ring", nside)
    np.testing.assert_array_equal(np.sort(a), np.arange(n))
    np.testing.assert_array_equal(b[a], np.arange(n))


def test_nside1_is_identity():
    for kind in ("ring2nest", "nest2ring"):
        np.testing.assert_array_equal(R.sigma(kind, 1), np.arange(12))


def test_ring_order_matches_ring_geometry():
    """RING index of a NEST pixel lies on the ring ix+iy predicts (the
    latitude index jr = (face//4 + 2) nside - ix - iy - 1)."""
    from oracle.geometry import healpix_ring_of_pixel
    nside = 8
    q = np.arange(12 * nside * nside)
    ix, iy, face = R.nest2xyf(q, nside)
    jr = ((face >> 2) + 2) * nside - ix - iy - 1
    ring = healpix_ring_of_pixel(nside)[R.sigma("nest2ring", nside)]
    np.testing.assert_array_equal(ring, jr - 1)


@pytest.mark.parametrize("nside", [1, 2, 8])
@pytest.mark.parametrize("origin", ORIGINS)
@pytest.mark.parametrize("cw", [False, True])
def test_flat_inverse_pairs(nside, origin, cw):
    n = np.arange(12 * nside * nside)
    n2f = R.sigma("nest2flat", nside, origin, cw)
    f2n = R.sigma("flat2nest", nside, origin, cw)
    r2f = R.sigma("ring2flat", nside, origin, cw)
    f2r = R.sigma("flat2ring", nside, origin, cw)
    np.testing.assert_array_equal(f2n[n2f], n)
    np.testing.assert_array_equal(f2r[r2f], n)
    # ring2flat = nest2flat after ring2nest
    np.testing.assert_array_equal(r2f, n2f[R.sigma("ring2nest", nside)])


def test_flat_south_ccw_is_row_major_morton_decode():
    nside = 8
    q = np.arange(12 * nside * nside)
    ix, iy, face = R.nest2xyf(q, nside)
    np.testing.assert_array_equal(R.sigma("nest2flat", nside, "S", False),
                                  face * nside * nside + iy * nside + ix)


@pytest.mark.parametrize("nside", [2, 8])
def test_flat2flat_identities(nside):
    n = np.arange(12 * nside * nside)
    for o in ORIGINS:
        for cw in (False, True):
            np.testing.assert_array_equal(R.sigma("flat2flat", nside, o, cw, o, cw), n)
            for o2 in ORIGINS:
                for cw2 in (False, True):
                    f2f = R.sigma("flat2flat", nside, o, cw, o2, cw2)
                    comp = R.sigma("nest2flat", nside, o2, cw2)[R.sigma("flat2nest", nside, o, cw)]
                    np.testing.assert_array_equal(f2f, comp)


def test_kernel_decomposition():
    """Every kind equals dst_scheme(q) o src_scheme(q)^-1 over NEST pixels q,
    which is how the tiled kernel evaluates it."""
    for nside in (1, 4, 16):
        q = np.arange(12 * nside * nside)
        ring = R.sigma("nest2ring", nside)
        assert np.array_equal(R.sigma("ring2nest", nside)[ring], q)
        for o in ORIGINS:
            for cw in (False, True):
                fl = R.sigma("nest2flat", nside, o, cw)
                assert np.array_equal(R.sigma("flat2nest", nside, o, cw)[fl], q)
                assert np.array_equal(R.sigma("ring2flat", nside, o, cw)[ring], fl)
                assert np.array_equal(R.sigma("flat2ring", nside, o, cw)[fl], ring)


def test_bad_origin_raises():
    with pytest.raises(RuntimeError):
        R.sigma("nest2flat", 2, "X", False)


def test_apply_scatter_semantics():
    sig = np.array([2, 0, 1])
    np.testing.assert_array_equal(R.apply(np.array([10, 20, 30]), sig), [20, 30, 10])

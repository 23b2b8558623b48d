"""Normalised associated Legendre functions P_lm(cos theta) in FP64 (oracle).

Test infrastructure only (see ``oracle/__init__.py``).

``legpoly`` restates cuhpx/sht_tools.py:232-256 (``legpoly_torch``, the
function SHTCUDA / iSHTCUDA use through ``_precompute_legpoly_torch``,
sht_tools.py:263-264) and sht_tools.py:192-229 (``legpoly``, used by the CPU
SHT / iSHT with norm="ortho"; its ``csphase`` branch multiplies by 1 and is a
no-op).  The functions are orthonormal on the sphere, without the
Condon-Shortley phase:

    P_00 = 1/sqrt(4 pi)
    P_ll = sqrt((2l+1)/(2l)) sin(theta) P_{l-1,l-1}
    P_{l,l-1}... written per m below:
    P_{m+1,m} = sqrt(2m+3) x P_mm
    P_lm = x a_lm P_{l-1,m} - b_lm P_{l-2,m},
        a_lm = sqrt((2l-1)(2l+1) / ((l-m)(l+m)))
        b_lm = sqrt((l+m-1)(l-m-1)(2l+1) / ((l-m)(l+m)(2l-3)))

The array layout follows the reference: ``out[m, l, k]`` for m < mmax,
l < lmax, zero for l < m.

``legpoly`` reproduces the reference's unscaled recurrence (values below the
FP64 range flush to zero, as in the reference).  ``legpoly_scaled`` runs the
same recurrence per m with a power-of-two exponent carried beside the value,
so values that the plain recurrence flushes are kept until they re-enter the
FP64 range; it is what the device table generator implements.
"""

import numpy as np


def recurrence_coeffs(l, m):
    """(a_lm, b_lm) of the three-term recurrence (sht_tools.py:250-251)."""
    l = np.asarray(l, dtype=np.float64)
    m = np.asarray(m, dtype=np.float64)
    a = np.sqrt((2 * l - 1) / (l - m) * (2 * l + 1) / (l + m))
    b = np.sqrt((l + m - 1) / (l - m) * (2 * l + 1) / (2 * l - 3) * (l - m - 1) / (l + m))
    return a, b


def legpoly(mmax, lmax, x):
    """Reference recurrence, vectorised over m and x; sht_tools.py:232-256."""
    x = np.asarray(x, dtype=np.float64)
    nmax = max(mmax, lmax)
    vdm = np.zeros((nmax, nmax, len(x)), dtype=np.float64)
    vdm[0, 0, :] = 1.0 / np.sqrt(4 * np.pi)
    for l in range(1, nmax):
        vdm[l - 1, l, :] = np.sqrt(2 * l + 1) * x * vdm[l - 1, l - 1, :]
        vdm[l, l, :] = np.sqrt((2 * l + 1) * (1 + x) * (1 - x) / 2 / l) * vdm[l - 1, l - 1, :]
    for l in range(2, nmax):
        m = np.arange(0, l - 1)
        a, b = recurrence_coeffs(l, m)
        vdm[m, l, :] = (x[None, :] * a[:, None] * vdm[m, l - 1, :]
                        - b[:, None] * vdm[m, l - 2, :])
    return vdm[:mmax, :lmax, :]


_SCALE_BITS = 600
_BIG = 2.0 ** _SCALE_BITS
_TINY = 2.0 ** -_SCALE_BITS


def legpoly_scaled_lrows(mmax, lmax, x):
    """Generator of (l, Q) with Q[m, k] = P_lm(x_k) for m < min(l + 1, mmax):
    the values of legpoly_scaled with l as the outer loop and all m advanced
    together (large vectors, one l-row in memory).  Same scaled recurrence:
    the sectoral start P_ll is the running product over l, each value carried
    as v * 2**(600 e) and emitted when e == 0."""
    x = np.asarray(x, dtype=np.float64)
    nx = len(x)
    s = np.sqrt((1 + x) * (1 - x))
    pmm = np.full(nx, 1.0 / np.sqrt(4 * np.pi))
    emm = np.zeros(nx, dtype=np.int64)
    cur = np.zeros((mmax, nx))                 # P_{l-1, m} (scaled)
    prev = np.zeros((mmax, nx))                # P_{l-2, m} (scaled)
    e = np.zeros((mmax, nx), dtype=np.int64)   # exponents of cur / prev
    scaled = np.zeros(mmax, dtype=bool)        # rows holding a negative exponent

    def rescale(r0, r1):
        rows = r0 + np.flatnonzero(scaled[r0:r1])
        if len(rows) == 0:
            return
        c, p, ee = cur[rows], prev[rows], e[rows]
        big = (ee < 0) & (np.abs(c) > 1.0)
        cur[rows] = np.where(big, c * _TINY, c)
        prev[rows] = np.where(big, p * _TINY, p)
        ee += big.astype(np.int64)
        e[rows] = ee
        scaled[rows] = (ee < 0).any(axis=1)

    for l in range(lmax):
        nm = min(l + 1, mmax)
        r = min(l - 1, mmax)                            # rows m <= l - 2
        if r > 0:
            a, b = recurrence_coeffs(l, np.arange(r))
            nxt = (x[None, :] * a[:, None]) * cur[:r] - b[:, None] * prev[:r]
            prev[:r] = cur[:r]
            cur[:r] = nxt
            rescale(0, r)
        if 1 <= l <= mmax:                              # m = l - 1: P_{m+1, m}
            m = l - 1
            prev[m] = cur[m]
            cur[m] = np.sqrt(2 * m + 3) * x * prev[m]
            rescale(m, m + 1)
        if l < mmax:                                    # m = l: sectoral
            if l > 0:
                pmm = pmm * np.sqrt((2 * l + 1) / (2 * l)) * s
                small = (np.abs(pmm) < _TINY) & (pmm != 0)
                pmm = np.where(small, pmm * _BIG, pmm)
                emm = emm - small.astype(np.int64)
            cur[l] = pmm
            prev[l] = 0.0
            e[l] = emm
            scaled[l] = (emm < 0).any()
        yield l, np.where(e[:nm] == 0, cur[:nm], 0.0)


def legpoly_scaled(mmax, lmax, x):
    """Same functions, recurrence in l per m with an exponent carried beside.

    Each value is held as v * 2**(600*e).  The sectoral start P_mm is built
    from P_00 by m multiplications, rescaling by 2**600 whenever |v| drops
    below 2**-600; the l-recurrence rescales both carried values together.
    A value is emitted as v * 2**(600*e) when e == 0 (and as zero when e < 0,
    i.e. when it is below ~1e-180 and so far below anything FP32 can hold).
    """
    x = np.asarray(x, dtype=np.float64)
    s = np.sqrt((1 + x) * (1 - x))
    out = np.zeros((mmax, lmax, len(x)), dtype=np.float64)
    pmm = np.full(len(x), 1.0 / np.sqrt(4 * np.pi))
    emm = np.zeros(len(x), dtype=np.int64)
    for m in range(mmax):
        if m > 0:
            pmm = pmm * np.sqrt((2 * m + 1) / (2 * m)) * s
            small = (np.abs(pmm) < _TINY) & (pmm != 0)
            pmm = np.where(small, pmm * _BIG, pmm)
            emm = emm - small.astype(np.int64)
        if m >= lmax:
            continue
        p2 = pmm.copy()
        e = emm.copy()
        out[m, m, :] = np.where(e == 0, p2, 0.0)
        if m + 1 >= lmax:
            continue
        p1 = np.sqrt(2 * m + 3) * x * p2
        big = (np.abs(p1) > 1.0) & (e < 0)
        p1 = np.where(big, p1 * _TINY, p1)
        p2 = np.where(big, p2 * _TINY, p2)
        e = e + big.astype(np.int64)
        out[m, m + 1, :] = np.where(e == 0, p1, 0.0)
        for l in range(m + 2, lmax):
            a, b = recurrence_coeffs(l, m)
            p0 = x * a * p1 - b * p2
            p2, p1 = p1, p0
            big = (np.abs(p1) > 1.0) & (e < 0)
            p1 = np.where(big, p1 * _TINY, p1)
            p2 = np.where(big, p2 * _TINY, p2)
            e = e + big.astype(np.int64)
            out[m, l, :] = np.where(e == 0, p1, 0.0)
    return out

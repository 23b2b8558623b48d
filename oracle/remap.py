"""RING / NESTED / flat (XY) pixel-index permutations (integer oracle).

Test infrastructure only (see ``oracle/__init__.py``).

Restates the index kernels of src/data_remapping/hpx_remapping_cuda.cu in
vectorised NumPy (int64 arithmetic, same formulas):
  * ring2xyf            hpx_remapping_cuda.cu:161-224
  * xyf2ring            hpx_remapping_cuda.cu:228-282
  * nest2xyf / xyf2nest hpx_remapping_cuda.cu:285-321 (bit spread 327-347)
  * correct_xy_orient   hpx_remapping_cuda.cu:100-134
  * flat <-> xyf        hpx_remapping_cuda.cu:33-97
and the wrappers' composition and scatter semantics ``out[sigma(i)] = in[i]``
(hpx_remapping_cuda.cu:137-146, 402-643, 902-917), including how each wrapper
derives ``flip`` and the rotation count ``k`` from the origins and the
clockwise flags (e.g. 415-418) and the origin codes of hpx_remapping.cpp:65-75
(S=0, E=1, N=2, W=3).

``sigma_*`` functions return, for every source index i, the destination index
sigma(i); ``apply`` performs the scatter on the last axis.
"""

import numpy as np

ORIGINS = {"S": 0, "E": 1, "N": 2, "W": 3}


def order_of(nside):
    return int(nside).bit_length() - 1


def _isqrt(v):
    """d_isqrt (hpx_remapping_cuda.cu:376-388) on int64 arrays."""
    v = np.asarray(v, dtype=np.int64)
    res = np.sqrt(v.astype(np.float64) + 0.5).astype(np.int64)
    res = np.where(res * res > v, res - 1, res)
    res = np.where((res + 1) * (res + 1) <= v, res + 1, res)
    return res


def spread_bits(v):
    v = np.asarray(v, dtype=np.int64) & 0xFFFF
    v = (v ^ (v << 8)) & 0x00FF00FF
    v = (v ^ (v << 4)) & 0x0F0F0F0F
    v = (v ^ (v << 2)) & 0x33333333
    v = (v ^ (v << 1)) & 0x55555555
    return v


def compress_bits(v):
    v = np.asarray(v, dtype=np.int64) & 0x55555555
    v = (v ^ (v >> 1)) & 0x33333333
    v = (v ^ (v >> 2)) & 0x0F0F0F0F
    v = (v ^ (v >> 4)) & 0x00FF00FF
    v = (v ^ (v >> 8)) & 0x0000FFFF
    return v


def ring2xyf(pix, nside):
    """RING index -> (ix, iy, face); hpx_remapping_cuda.cu:161-224."""
    pix = np.asarray(pix, dtype=np.int64)
    order = order_of(nside)
    npface = nside * nside
    ncap = (npface - nside) << 1
    npix = 12 * npface
    iring = np.zeros_like(pix)
    iphi = np.zeros_like(pix)
    kshift = np.zeros_like(pix)
    nr = np.zeros_like(pix)
    face = np.zeros_like(pix)

    north = pix < ncap
    south = pix >= npix - ncap
    equ = ~north & ~south

    p = pix[north]
    ir = (1 + _isqrt(1 + 2 * p)) >> 1
    ip = (p + 1) - 2 * ir * (ir - 1)
    iring[north] = ir
    iphi[north] = ip
    nr[north] = ir
    face[north] = (ip - 1) // ir

    p = pix[equ]
    ipq = p - ncap
    tmp = ipq >> (order + 2)
    ir = tmp + nside
    ip = ipq - tmp * 4 * nside + 1
    iring[equ] = ir
    iphi[equ] = ip
    kshift[equ] = (ir + nside) & 1
    nr[equ] = nside
    ire = tmp + 1
    irm = 2 * nside + 1 - tmp
    ifm = (ip - (ire >> 1) + nside - 1) >> order
    ifp = (ip - (irm >> 1) + nside - 1) >> order
    face[equ] = np.where(ifp == ifm, ifp | 4, np.where(ifp < ifm, ifp, ifm + 8))

    p = pix[south]
    ipq = npix - p
    ir = (1 + _isqrt(2 * ipq - 1)) >> 1
    ip = 4 * ir + 1 - (ipq - 2 * ir * (ir - 1))
    nr[south] = ir
    iring[south] = 4 * nside - ir
    iphi[south] = ip
    face[south] = (ip - 1) // ir + 8

    irt = iring - ((2 + (face >> 2)) * nside) + 1
    ipt = ((face & 3) << 1) - ((face >> 2) & 1) + 1
    ipt = 2 * iphi - ipt * nr - kshift - 1
    ipt = np.where(ipt >= 2 * nside, ipt - 8 * nside, ipt)
    ix = (ipt - irt) >> 1
    iy = (-ipt - irt) >> 1
    return ix, iy, face


def xyf2ring(ix, iy, face, nside):
    """(ix, iy, face) -> RING index; hpx_remapping_cuda.cu:228-282."""
    ix = np.asarray(ix, dtype=np.int64)
    iy = np.asarray(iy, dtype=np.int64)
    face = np.asarray(face, dtype=np.int64)
    nl4 = 4 * nside
    npface = nside * nside
    ncap = (npface - nside) << 1
    npix = 12 * npface
    jr = ((face >> 2) + 2) * nside - ix - iy - 1
    north = jr < nside
    south = jr >= 3 * nside
    equ = ~north & ~south
    ring_nr = 4 * nside - jr
    nr = np.where(north, 4 * jr, np.where(equ, 4 * nside, 4 * ring_nr))
    n_before = np.where(north, 2 * jr * (jr - 1),
                        np.where(equ, ncap + (jr - nside) * 4 * nside,
                                 npix - 2 * ring_nr * (ring_nr + 1)))
    shifted = np.where(equ, ((jr - nside) & 1) == 0, True)
    nr = nr >> 2
    kshift = 1 - shifted.astype(np.int64)
    jp = ((face & 3) << 1) - ((face >> 2) & 1) + 1
    # C integer division truncates toward zero; the numerator is >= 0 here
    # except where the wrap below applies, so use explicit truncation.
    num = jp * nr + ix - iy + 1 + kshift
    jp = np.where(num >= 0, num // 2, -((-num) // 2))
    jp = np.where(jp < 1, jp + nl4, jp)
    return n_before + jp - 1


def nest2xyf(pix, nside):
    """NEST index -> (ix, iy, face); hpx_remapping_cuda.cu:285-304."""
    pix = np.asarray(pix, dtype=np.int64)
    order = order_of(nside)
    face = pix >> (2 * order)
    p = pix & (nside * nside - 1)
    return compress_bits(p), compress_bits(p >> 1), face


def xyf2nest(ix, iy, face, nside):
    """(ix, iy, face) -> NEST index; hpx_remapping_cuda.cu:307-321."""
    order = order_of(nside)
    return (np.asarray(face, dtype=np.int64) << (2 * order)) + spread_bits(ix) + (spread_bits(iy) << 1)


def correct_xy_orient(ix, iy, nside, flip, k):
    """Optional ix<->iy swap, then k quarter turns; hpx_remapping_cuda.cu:100-134."""
    ix = np.asarray(ix, dtype=np.int64)
    iy = np.asarray(iy, dtype=np.int64)
    if flip:
        ix, iy = iy, ix
    if k == 1:
        nx, ny = -iy - 1, ix
    elif k == 2:
        nx, ny = -ix - 1, -iy - 1
    elif k == 3:
        nx, ny = iy, -ix - 1
    else:
        nx, ny = ix, iy
    return nx & (nside - 1), ny & (nside - 1)


def xyf2flat(ix, iy, face, nside, flip, k):
    """hpx_remapping_cuda.cu:55-72."""
    order = order_of(nside)
    ix, iy = correct_xy_orient(ix, iy, nside, flip, k)
    return (np.asarray(face, dtype=np.int64) << (2 * order)) + (iy << order) + ix


def flat2xyf(pix, nside, flip, k):
    """hpx_remapping_cuda.cu:76-97."""
    pix = np.asarray(pix, dtype=np.int64)
    order = order_of(nside)
    face = pix >> (2 * order)
    iy = (pix >> order) & (nside - 1)
    ix = pix & (nside - 1)
    ix, iy = correct_xy_orient(ix, iy, nside, flip, k)
    return ix, iy, face


def _flip_k(src_origin, src_cw, dst_origin, dst_cw):
    """Wrapper parameters (e.g. hpx_remapping_cuda.cu:415-418)."""
    flip = bool(src_cw) != bool(dst_cw)
    rotations = dst_origin - src_origin
    k = (-rotations if dst_cw else rotations) & 3
    return flip, k


def _origin(o):
    if isinstance(o, str):
        if o not in ORIGINS:
            raise RuntimeError("The origin must be one from S, E, N, W. Stop")
        return ORIGINS[o]
    return int(o)


def sigma(kind, nside, origin="S", clockwise=False, dst_origin="S", dst_clockwise=False):
    """Destination index of every source index for one remap kind.

    kind: ring2nest, nest2ring, nest2flat, flat2nest, ring2flat, flat2ring,
    flat2flat.  For the one-sided kinds (origin, clockwise) describe the flat
    layout; for flat2flat they are the source layout and (dst_origin,
    dst_clockwise) the destination layout.
    """
    npix = 12 * nside * nside
    pix = np.arange(npix, dtype=np.int64)
    o = _origin(origin)
    if kind == "ring2nest":
        return xyf2nest(*ring2xyf(pix, nside), nside)
    if kind == "nest2ring":
        return xyf2ring(*nest2xyf(pix, nside), nside)
    if kind == "nest2flat":  # nest2xy: src (0, ccw) -> dst (o, cw)
        flip, k = _flip_k(0, False, o, clockwise)
        ix, iy, f = nest2xyf(pix, nside)
        return xyf2flat(ix, iy, f, nside, flip, k)
    if kind == "flat2nest":  # xy2nest: src (o, cw) -> dst (0, ccw)
        flip, k = _flip_k(o, clockwise, 0, False)
        return xyf2nest(*flat2xyf(pix, nside, flip, k), nside)
    if kind == "ring2flat":
        flip, k = _flip_k(0, False, o, clockwise)
        ix, iy, f = ring2xyf(pix, nside)
        return xyf2flat(ix, iy, f, nside, flip, k)
    if kind == "flat2ring":
        flip, k = _flip_k(o, clockwise, 0, False)
        return xyf2ring(*flat2xyf(pix, nside, flip, k), nside)
    if kind == "flat2flat":
        d = _origin(dst_origin)
        flip, k = _flip_k(o, clockwise, d, dst_clockwise)
        order = order_of(nside)
        face = pix >> (2 * order)
        iy = (pix >> order) & (nside - 1)
        ix = pix & (nside - 1)
        ix, iy = correct_xy_orient(ix, iy, nside, flip, k)
        return (face << (2 * order)) + (iy << order) + ix
    raise ValueError(f"unknown remap kind {kind}")


def apply(data, sig):
    """out[..., sig[i]] = data[..., i] (the reference's scatter)."""
    data = np.asarray(data)
    out = np.empty_like(data)
    out[..., sig] = data
    return out

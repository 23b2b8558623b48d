"""RING / NESTED / flat (XY) pixel remaps on the GPU.

Drop-in for cuhpx/hpx_remap.py:20-102 (``ring2nest``, ``nest2ring``,
``nest2flat``, ``flat2nest``, ``ring2flat``, ``flat2ring``, ``flat2flat``),
with the argument order, the errors and the output allocation of the
reference: a non-power-of-two nside raises ValueError (hpx_remap.py:26-27), a
last dimension other than 12*nside^2 raises RuntimeError
(hpx_remapping.cpp:25), an origin other than S/E/N/W raises RuntimeError
(hpx_remapping.cpp:74), and the result is a new tensor shaped like the input.
All leading dimensions are flattened into rows, as the reference's _batch
entry points do (hpx_remapping_cuda.cu:1247-1293).

Every call is one launch of the tiled permutation kernel in libhpxmi.so on
the current stream; it is bit-exact for any dtype whose element is 1, 2, 4 or
8 bytes.
"""

import torch

from . import _lib

_ORIGINS = {"S": 0, "E": 1, "N": 2, "W": 3}


def is_power_of_two(n):
    return (n > 0) and (n & (n - 1)) == 0


def _origin(o, what="origin"):
    if o not in _ORIGINS:
        raise RuntimeError(f"The {what} must be one from S, E, N, W. Stop")
    return _ORIGINS[o]


def _remap(kind, data, nside, origin="S", clockwise=False, dst_origin="S", dst_clockwise=False,
           what="origin"):
    if not is_power_of_two(nside):
        raise ValueError("nside must be a power of two")
    if data.dim() == 0:
        raise RuntimeError("data must have at least one dimension")
    npix = 12 * nside * nside
    if data.size(-1) != npix:
        raise RuntimeError(
            "The number of elements in the input array is not equal to the number of HEALPix "
            "grid at current nside. Stop.")
    if kind == "flat2flat":
        o = _origin(origin, "src origin")
        d = _origin(dst_origin, "dest origin")
    else:
        o = _origin(origin, what)
        d = 0
    if not data.is_cuda:
        raise RuntimeError("cuhpx_amd remaps need a GPU tensor")
    esize = data.element_size()
    if esize not in (1, 2, 4, 8):
        raise TypeError(f"unsupported element size {esize} ({data.dtype})")
    src = data.contiguous()
    out = torch.empty_like(src)
    nrows = src.numel() // npix
    with torch.cuda.device(src.device):
        rc = _lib.lib().hpx_remap(_lib.REMAP_KINDS[kind], nside, o, int(bool(clockwise)), d,
                                  int(bool(dst_clockwise)), _lib.ptr(src), _lib.ptr(out), nrows,
                                  esize, _lib.stream_ptr(src.device))
    _lib.check(rc, kind)
    return out


def ring2nest(data: torch.Tensor, nside: int):
    """RING -> NESTED (hpx_remap.py:24-32)."""
    return _remap("ring2nest", data, nside)


def nest2ring(data: torch.Tensor, nside: int):
    """NESTED -> RING (hpx_remap.py:35-43)."""
    return _remap("nest2ring", data, nside)


def nest2flat(data: torch.Tensor, origin: str, clockwise: bool, nside: int):
    """NESTED -> flat XY layout with the given origin and direction (hpx_remap.py:46-54)."""
    return _remap("nest2flat", data, nside, origin, clockwise)


def flat2nest(data: torch.Tensor, origin: str, clockwise: bool, nside: int):
    """flat XY layout -> NESTED (hpx_remap.py:57-65)."""
    return _remap("flat2nest", data, nside, origin, clockwise)


def ring2flat(data: torch.Tensor, origin: str, clockwise: bool, nside: int):
    """RING -> flat XY layout (hpx_remap.py:68-76)."""
    return _remap("ring2flat", data, nside, origin, clockwise)


def flat2ring(data: torch.Tensor, origin: str, clockwise: bool, nside: int):
    """flat XY layout -> RING (hpx_remap.py:79-87)."""
    return _remap("flat2ring", data, nside, origin, clockwise)


def flat2flat(data: torch.Tensor, src_origin: str, src_clockwise: bool, dest_origin: str,
              dest_clockwise: bool, nside: int):
    """flat XY layout -> flat XY layout of another origin / direction (hpx_remap.py:90-102)."""
    return _remap("flat2flat", data, nside, src_origin, src_clockwise, dest_origin, dest_clockwise)

"""Remap kernels vs the integer oracle: bit-exact, every kind, origin and
direction, several dtypes, ragged batches, and the full nside=1024 x 64 case."""

import numpy as np
import pytest
import torch

import cuhpx_amd
from oracle import remap as R

pytestmark = pytest.mark.gpu

ORIGINS = ["S", "E", "N", "W"]
ONE_SIDED = ["nest2flat", "flat2nest", "ring2flat", "flat2ring"]


def _expected(kind, data_np, nside, *args):
    return R.apply(data_np, R.sigma(kind, nside, *args))


def _call(kind, x, nside, *args):
    if kind in ("ring2nest", "nest2ring"):
        return getattr(cuhpx_amd, kind)(x, nside)
    if kind == "flat2flat":
        o, cw, o2, cw2 = args
        return cuhpx_amd.flat2flat(x, o, cw, o2, cw2, nside)
    o, cw = args
    return getattr(cuhpx_amd, kind)(x, o, cw, nside)


def _payload(nside, rows):
    npix = 12 * nside * nside
    return (torch.arange(rows * npix, dtype=torch.int64) * 7 + 3).reshape(rows, npix)


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 64, 128])
def test_ring_nest_all_rows(nside):
    x = _payload(nside, 3)
    for kind in ("ring2nest", "nest2ring"):
        got = _call(kind, x.cuda(), nside).cpu().numpy()
        np.testing.assert_array_equal(got, _expected(kind, x.numpy(), nside), err_msg=f"{kind} {nside}")


@pytest.mark.parametrize("nside", [1, 4, 32, 128])
@pytest.mark.parametrize("kind", ONE_SIDED)
def test_flat_kinds_all_origins(nside, kind):
    x = _payload(nside, 2)
    xd = x.cuda()
    for o in ORIGINS:
        for cw in (False, True):
            got = _call(kind, xd, nside, o, cw).cpu().numpy()
            np.testing.assert_array_equal(got, _expected(kind, x.numpy(), nside, o, cw),
                                          err_msg=f"{kind} nside={nside} origin={o} cw={cw}")


@pytest.mark.parametrize("nside", [2, 16, 128])
def test_flat2flat_all_pairs(nside):
    x = _payload(nside, 2)
    xd = x.cuda()
    for o in ORIGINS:
        for cw in (False, True):
            for o2 in ORIGINS:
                for cw2 in (False, True):
                    got = _call("flat2flat", xd, nside, o, cw, o2, cw2).cpu().numpy()
                    exp = _expected("flat2flat", x.numpy(), nside, o, cw, o2, cw2)
                    np.testing.assert_array_equal(got, exp, err_msg=f"{nside} {o}{cw}->{o2}{cw2}")


@pytest.mark.parametrize("dtype", [torch.uint8, torch.int8, torch.int16, torch.int32, torch.int64,
                                   torch.float32, torch.float64, torch.bfloat16, torch.float16,
                                   torch.bool, torch.complex64])
def test_dtypes_are_bit_exact(dtype):
    nside = 16
    npix = 12 * nside * nside
    g = torch.Generator().manual_seed(7)
    if dtype.is_floating_point or dtype.is_complex:
        x = torch.randn(2, npix, generator=g).to(dtype)
        if dtype.is_complex:
            x = x + 1j * torch.randn(2, npix, generator=g).to(dtype)
    elif dtype == torch.bool:
        x = torch.rand(2, npix, generator=g) > 0.5
    else:
        x = torch.randint(-100, 100, (2, npix), generator=g).to(dtype) if dtype != torch.uint8 else \
            torch.randint(0, 255, (2, npix), generator=g).to(dtype)
    for kind, args in [("nest2ring", ()), ("ring2nest", ()), ("nest2flat", ("W", True)),
                       ("flat2flat", ("E", False, "N", True))]:
        sig = torch.from_numpy(R.sigma(kind, nside, *args))
        exp = torch.empty_like(x)
        exp[:, sig] = x
        got = _call(kind, x.cuda(), nside, *args).cpu()
        assert torch.equal(got.view(torch.uint8) if dtype == torch.bool else got,
                           exp.view(torch.uint8) if dtype == torch.bool else exp), (kind, dtype)


def test_one_dim_and_leading_dims_match_row_loop():
    nside = 8
    npix = 12 * nside * nside
    x = torch.randn(2, 3, npix, device="cuda")
    batched = cuhpx_amd.ring2nest(x, nside)
    assert batched.shape == x.shape
    for i in range(2):
        for j in range(3):
            assert torch.equal(batched[i, j], cuhpx_amd.ring2nest(x[i, j], nside))


def test_non_contiguous_input():
    nside = 4
    npix = 12 * nside * nside
    x = torch.randn(npix, 3, device="cuda").t()  # non-contiguous rows
    got = cuhpx_amd.nest2ring(x, nside)
    assert torch.equal(got, cuhpx_amd.nest2ring(x.contiguous(), nside))


def test_round_trips_are_identity():
    nside = 32
    x = torch.randn(4, 12 * nside * nside, device="cuda")
    assert torch.equal(cuhpx_amd.nest2ring(cuhpx_amd.ring2nest(x, nside), nside), x)
    for o in ORIGINS:
        for cw in (False, True):
            assert torch.equal(cuhpx_amd.flat2nest(cuhpx_amd.nest2flat(x, o, cw, nside), o, cw, nside), x)
            assert torch.equal(cuhpx_amd.flat2ring(cuhpx_amd.ring2flat(x, o, cw, nside), o, cw, nside), x)


@pytest.mark.parametrize("kind,args", [("nest2ring", ()), ("ring2nest", ()), ("nest2flat", ("N", True)),
                                       ("flat2flat", ("S", False, "W", True))])
def test_full_size_nside1024_batch64(kind, args):
    """The measured configuration: [64, 12*1024^2] float32 plus an int32 tensor
    carrying each pixel's own index.  out_idx[b, sigma(i)] = i, so every row of
    the index output must equal sigma^-1, and the float output must be the
    float input gathered through it."""
    nside, rows = 1024, 64
    npix = 12 * nside * nside
    print(f"[full-size] {kind} {args}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(2)
    xf = torch.randn(rows, npix, device="cuda", generator=g)
    xi = torch.arange(npix, dtype=torch.int32, device="cuda").expand(rows, npix).contiguous()
    sig = R.sigma(kind, nside, *args)
    inv = np.empty_like(sig)
    inv[sig] = np.arange(npix)
    inv_t = torch.from_numpy(inv.astype(np.int32)).cuda()
    oi = _call(kind, xi, nside, *args)
    assert torch.equal(oi, inv_t.expand(rows, npix))
    of = _call(kind, xf, nside, *args)
    assert torch.equal(of, torch.gather(xf, 1, oi.long()))


@pytest.mark.parametrize("nside,rows,dtype", [(512, 3, torch.int32), (512, 2, torch.int16),
                                              (512, 2, torch.uint8), (1024, 2, torch.int16),
                                              (1024, 2, torch.uint8), (2048, 2, torch.int32)])
def test_nest2ring_band_tiles(nside, rows, dtype):
    """nest2ring at nside >= 512 with pixels of at most 4 bytes walks the map in
    parallelogram tiles that follow the rings (32 rings x 512 pixels from
    nside 1024, 64 x 256 at nside 512, partly filled at each face's corners):
    bit-exact against the oracle for every element size of both tile shapes,
    and at one nside above the measured size."""
    print(f"[band tiles] nest2ring nside={nside} rows={rows} {dtype}", flush=True)
    x = _payload(nside, rows).to(dtype)
    got = cuhpx_amd.nest2ring(x.cuda(), nside).cpu().numpy()
    np.testing.assert_array_equal(got, _expected("nest2ring", x.numpy(), nside))


@pytest.mark.parametrize("nside,rows,dtype", [(64, 1, torch.int32), (64, 7, torch.int32), (128, 5, torch.int32),
                                              (256, 6, torch.int32), (64, 3, torch.uint8), (128, 5, torch.int16),
                                              (128, 5, torch.int64)])
def test_nest_flat_split_tiles_ragged_rows(nside, rows, dtype):
    """NEST <-> flat and flat -> flat from nside 64 run 64 x 64 tiles whose
    loads and stores are done by separate halves of the work-group, four rows
    per work-group: row counts that leave a short last chunk, 1-, 2-, 4- and
    8-byte pixels, all four origins both ways, and float32 bit patterns,
    against the oracle."""
    x = _payload(nside, rows).to(dtype)
    xd = x.cuda()
    for o in ORIGINS:
        for cw in (False, True):
            print(f"[split tiles] nside={nside} rows={rows} {dtype} origin={o} cw={cw}", flush=True)
            for kind in ("nest2flat", "flat2nest"):
                got = _call(kind, xd, nside, o, cw).cpu().numpy()
                np.testing.assert_array_equal(got, _expected(kind, x.numpy(), nside, o, cw),
                                              err_msg=f"{kind} nside={nside} rows={rows} {o}{cw}")
            o2, cw2 = ORIGINS[(ORIGINS.index(o) + 1) % 4], not cw
            got = _call("flat2flat", xd, nside, o, cw, o2, cw2).cpu().numpy()
            np.testing.assert_array_equal(got, _expected("flat2flat", x.numpy(), nside, o, cw, o2, cw2))
    xf = torch.randn(rows, 12 * nside * nside, device="cuda")
    got = cuhpx_amd.nest2flat(xf, "N", True, nside)
    exp = _expected("nest2flat", xf.view(torch.int32).cpu().numpy(), nside, "N", True)
    np.testing.assert_array_equal(got.view(torch.int32).cpu().numpy(), exp)


@pytest.mark.parametrize("nside,rows,dtype", [(128, 9, torch.int32), (256, 13, torch.int16), (128, 11, torch.uint8),
                                              (128, 9, torch.int64), (512, 3, torch.int64)])
def test_ring_kinds_split_tiles_ragged_rows(nside, rows, dtype):
    """The RING kinds from nside 128 run 64 x 64 tiles whose loads and stores
    are done by separate halves of the work-group, eight rows per work-group
    (except NEST -> RING at nside >= 512 with pixels of at most 4 bytes, and
    flat -> RING with 4-byte pixels there): row counts that leave a short
    last chunk, every element size, all four RING kinds, two flat origins."""
    x = _payload(nside, rows).to(dtype)
    xd = x.cuda()
    print(f"[ring split tiles] nside={nside} rows={rows} {dtype}", flush=True)
    for kind in ("ring2nest", "nest2ring"):
        got = _call(kind, xd, nside).cpu().numpy()
        np.testing.assert_array_equal(got, _expected(kind, x.numpy(), nside), err_msg=f"{kind} {nside} {rows}")
    for o, cw in (("N", True), ("E", False)):
        for kind in ("ring2flat", "flat2ring"):
            got = _call(kind, xd, nside, o, cw).cpu().numpy()
            np.testing.assert_array_equal(got, _expected(kind, x.numpy(), nside, o, cw),
                                          err_msg=f"{kind} {nside} {rows} {o}{cw}")


@pytest.mark.parametrize("nside", [512, 1024])
def test_flat2ring_large_nside_all_origins(nside):
    """flat2ring (a rotated / mirrored flat source, a RING destination) at the
    sizes where the RING-kind tiles are 128 pixels wide: bit-exact against the
    oracle for all four origins in both directions, plus 1- and 2-byte pixels
    for one orientation."""
    x = _payload(nside, 2)
    x32 = x.to(torch.int32)
    xd = x32.cuda()
    for o in ORIGINS:
        for cw in (False, True):
            print(f"[large nside] flat2ring nside={nside} origin={o} cw={cw}", flush=True)
            got = _call("flat2ring", xd, nside, o, cw).cpu().numpy()
            np.testing.assert_array_equal(got, _expected("flat2ring", x32.numpy(), nside, o, cw),
                                          err_msg=f"flat2ring nside={nside} origin={o} cw={cw}")
    for dtype in (torch.int16, torch.uint8):
        xs = x.to(dtype)
        got = _call("flat2ring", xs.cuda(), nside, "W", True).cpu().numpy()
        np.testing.assert_array_equal(got, _expected("flat2ring", xs.numpy(), nside, "W", True))


def test_full_size_nside1024_all_origins():
    """Origin coverage at the measured size ([64, 12*1024^2] int32 index rows):
    nest2flat for all four origins in both directions against the oracle
    permutation, then flat2flat for all 64 (source, destination) pairs, whose
    output must be the destination's flat layout of the same map:
    flat2flat(a -> b)(nest2flat_a(x)) = nest2flat_b(x)."""
    nside, rows = 1024, 64
    npix = 12 * nside * nside
    xi = torch.arange(npix, dtype=torch.int32, device="cuda").expand(rows, npix).contiguous()
    combos = [(o, cw) for o in "ESWN" for cw in (False, True)]
    flats = {}
    for o, cw in combos:
        print(f"[full-size] nest2flat {o} cw={cw}", flush=True)
        sig = R.sigma("nest2flat", nside, o, cw)
        inv = np.empty_like(sig)
        inv[sig] = np.arange(npix)
        out = cuhpx_amd.nest2flat(xi, o, cw, nside)
        assert torch.equal(out, torch.from_numpy(inv.astype(np.int32)).cuda().expand(rows, npix))
        flats[(o, cw)] = out
    print("[full-size] flat2flat, 64 pairs", flush=True)
    for (so, scw) in combos:
        for (do, dcw) in combos:
            got = cuhpx_amd.flat2flat(flats[(so, scw)], so, scw, do, dcw, nside)
            assert torch.equal(got, flats[(do, dcw)]), (so, scw, do, dcw)
            del got
    torch.cuda.synchronize()

"""Launch-shape variants (hpx_sht_set_shapes): every FFT and GEMM kernel
variant gives bit-identical results, and none reads workspace that an
earlier stage of the same call left unwritten (NaN-filled workspace)."""

import itertools

import pytest
import torch

from cuhpx_amd import _lib
from cuhpx_amd.hpx_sht import ShtPlan

pytestmark = pytest.mark.gpu

DEFAULT = {"fft_analysis_waves": [16, 4, 2], "fft_synthesis_waves": [16, 4, 2], "gemm_fwd_col_tiles": 1,
           "gemm_inv_col_tiles": 1, "gemm_inv_row_blocks": 1, "fft_analysis_threads_per_field": [64, 128, 64],
           "fft_synthesis_threads_per_field": [64, 128, 64]}
GEMM_VARIANTS = [(fwd, inv, rh) for fwd in (1, 2, 4) for inv in (1, 2) for rh in (1, 2)]


def _run_all(plan, n, f, a, prec, fill):
    """forward, inverse and both adjoints on a workspace filled with `fill`."""
    lib = _lib.lib()
    s = _lib.stream_ptr(torch.device("cuda"))
    L, M = plan.lmax, plan.mmax
    outs = []
    for fn, extra, inp, shape in [(lib.hpx_sht_forward, (prec,), f, (n, L, M, 2)),
                                  (lib.hpx_sht_inverse, (prec,), a, (n, plan.npix)),
                                  (lib.hpx_sht_forward_adjoint, (prec, 0), a, (n, plan.npix)),
                                  (lib.hpx_sht_inverse_adjoint, (prec, 0), f, (n, L, M, 2))]:
        ws, nbytes = plan.workspace(n)
        ws.view(torch.float32).fill_(fill)
        out = torch.zeros(shape, device="cuda")
        rc = fn(plan.handle, *extra, _lib.ptr(inp), _lib.ptr(out), n, _lib.ptr(ws), nbytes, s)
        assert rc == 0, lib.hpx_last_error()
        outs.append(out)
    torch.cuda.synchronize()
    return outs


@pytest.mark.parametrize("nside,L,n", [(32, 65, 40), (64, 129, 3)])
@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_gemm_variants_bit_identical_and_poison_free(nside, L, n, prec):
    plan = ShtPlan.healpix(nside, L, L)
    g = torch.Generator(device="cuda").manual_seed(11)
    f = torch.randn(n, plan.npix, device="cuda", generator=g)
    a = torch.randn(n, L, L, 2, device="cuda", generator=g)
    p = _lib.PRECISIONS[prec]
    ref = None
    for fwd, inv, rh in GEMM_VARIANTS:
        shapes = dict(DEFAULT, gemm_fwd_col_tiles=fwd, gemm_inv_col_tiles=inv, gemm_inv_row_blocks=rh)
        plan.set_shapes(prec, n, shapes)
        assert plan.tuned_shapes(prec, n) == shapes
        clean = _run_all(plan, n, f, a, p, 0.0)
        poisoned = _run_all(plan, n, f, a, p, float("nan"))
        print(f"[gemm nside={nside} n={n} {prec}] fwd {fwd} inv {inv}x{rh}", flush=True)
        for c, q in zip(clean, poisoned):
            assert torch.isfinite(c).all()
            assert torch.equal(c, q), (fwd, inv, rh)
        if ref is None:
            ref = clean
        for c, r in zip(clean, ref):
            assert torch.equal(c, r), (fwd, inv, rh)


@pytest.mark.parametrize("cls", [0, 1])
def test_fft_variants_bit_identical(cls):
    """nside 128: class 0 (rings up to 512 points and the 512-point equator
    pairs) and class 1 (1024-point Bluestein rings are class 0 at nside 128,
    the class-1 kernels serve 2048-point lengths: nside 256)."""
    nside = 128 if cls == 0 else 256
    L = 2 * nside + 1
    n = 5
    plan = ShtPlan.healpix(nside, L, L)
    g = torch.Generator(device="cuda").manual_seed(12)
    f = torch.randn(n, plan.npix, device="cuda", generator=g)
    a = torch.randn(n, L, L, 2, device="cuda", generator=g)
    p = _lib.PRECISIONS["bf16x3"]
    ref = None
    waves = [[4, 8, 16], [2, 4, 8]][cls]
    for w, tpf in itertools.product(waves, (64, 128, 256)):
        if w * tpf > 1024:
            continue
        aw = list(DEFAULT["fft_analysis_waves"])
        at = list(DEFAULT["fft_analysis_threads_per_field"])
        aw[cls], at[cls] = w, tpf
        shapes = dict(DEFAULT, fft_analysis_waves=aw, fft_synthesis_waves=aw, fft_analysis_threads_per_field=at,
                      fft_synthesis_threads_per_field=at)
        try:
            plan.set_shapes("bf16x3", n, shapes)
        except NotImplementedError as exc:  # a shape whose LDS exceeds the work-group's
            assert "LDS" in str(exc)
            continue
        out = _run_all(plan, n, f, a, p, float("nan"))
        print(f"[fft class {cls}] {w} fields x {tpf} threads", flush=True)
        if ref is None:
            ref = out
        for c, r in zip(out, ref):
            assert torch.equal(c, r), (w, tpf)


def test_set_shapes_refuses_unknown_variants():
    plan = ShtPlan.healpix(16, 33, 33)  # 32 ring pairs: no 8-wave forward GEMM
    for bad in [dict(DEFAULT, gemm_fwd_col_tiles=4), dict(DEFAULT, gemm_fwd_col_tiles=3),
                dict(DEFAULT, gemm_inv_row_blocks=3), dict(DEFAULT, fft_analysis_waves=[5, 4, 2]),
                dict(DEFAULT, fft_synthesis_threads_per_field=[64, 512, 64])]:
        with pytest.raises((ValueError, NotImplementedError)):
            plan.set_shapes("bf16x3", 4, bad)
    plan.set_shapes("bf16x3", 4, DEFAULT)
    assert plan.tuned_shapes("bf16x3", 4) == DEFAULT

"""Autotuned launch shapes and mixed-size requests through one HIP graph.

Launch shapes change only how work is split over work-groups, never the
arithmetic of an output element, so tuned and untuned runs and the graph and
per-size calls must agree bit for bit; the graph's outputs are also checked
against the FP64 oracle at the split-BF16 tolerance.
"""

import numpy as np
import pytest
import torch

import cuhpx_amd
from cuhpx_amd.hpx_sht import ShtPlan
from oracle import sht as S

pytestmark = pytest.mark.gpu


def rel_l2(x, ref):
    return float(np.linalg.norm(np.asarray(x) - ref) / np.linalg.norm(ref))


def band_limited_alm(lmax, mmax, nfield, seed):
    rng = np.random.default_rng(seed)
    l = np.arange(lmax)[:, None]
    cl = (l + 1.0) ** -2
    a = (rng.standard_normal((nfield, lmax, mmax)) + 1j * rng.standard_normal((nfield, lmax, mmax)))
    a *= np.sqrt(cl / 2)
    a[:, :, 0] = rng.standard_normal((nfield, lmax)) * np.sqrt(cl[:, 0])
    a[:, l[:, 0][:, None] < np.arange(mmax)[None, :]] = 0
    return a


@pytest.mark.parametrize("nside,nbatch", [(16, 3), (32, 40)])
def test_autotune_keeps_results_bit_identical(nside, nbatch):
    L = 2 * nside + 1
    plan = ShtPlan.healpix(nside, L, L)
    f = torch.randn(nbatch, plan.npix, device="cuda")
    a0 = plan.forward(f, "bf16x3")
    g0 = plan.inverse(a0, "bf16x3")
    shapes = plan.autotune("bf16x3", nbatch)
    print(f"[autotune nside={nside} batch={nbatch}] {shapes}", flush=True)
    assert all(w in (1, 2, 4, 8, 16) for w in shapes["fft_analysis_waves"] + shapes["fft_synthesis_waves"])
    # forward GEMM: 1 or 2 column tiles of 64, or 4 = the 8-wave kernel (128 columns)
    assert shapes["gemm_fwd_col_tiles"] in (1, 2, 4) and shapes["gemm_inv_col_tiles"] in (1, 2)
    assert shapes["gemm_inv_row_blocks"] in (1, 2)
    assert all(t in (64, 128, 256) for t in shapes["fft_analysis_threads_per_field"]
               + shapes["fft_synthesis_threads_per_field"])
    assert plan.tuned_shapes("bf16x3", nbatch) == shapes
    assert torch.equal(plan.forward(f, "bf16x3"), a0)
    assert torch.equal(plan.inverse(a0, "bf16x3"), g0)


SIZES = [(4, 3), (8, 2), (16, 5), (32, 1)]


def test_mixed_graph_matches_per_size_calls_and_oracle():
    specs = [(n, b, 2 * n + 1, 2 * n + 1) for n, b in SIZES]
    mixed = cuhpx_amd.MixedSHT(specs, precision="bf16x3", direction="both")
    maps64 = []
    for i, (n, b, L, _) in enumerate(specs):
        o = S.SHTOracle.healpix(n, L, L)
        maps64.append(o.inverse(band_limited_alm(L, L, b, 20 + i)))
    alm, back = mixed([torch.from_numpy(m.astype(np.float32)) for m in maps64])
    torch.cuda.synchronize()
    alm = [a.clone() for a in alm]
    back = [g.clone() for g in back]
    for i, (n, b, L, _) in enumerate(specs):
        sht = cuhpx_amd.SHTCUDA(n, lmax=L, mmax=L, precision="bf16x3")
        isht = cuhpx_amd.iSHTCUDA(n, lmax=L, mmax=L, precision="bf16x3")
        x = mixed.maps_in[i]
        assert torch.equal(alm[i], sht(x)), f"nside {n} forward differs from the per-size call"
        assert torch.equal(back[i], isht(alm[i])), f"nside {n} inverse differs from the per-size call"
        o = S.SHTOracle.healpix(n, L, L)
        x64 = x.double().cpu().numpy()
        err_a = rel_l2(alm[i].cpu().numpy(), o.forward(x64))
        err_f = rel_l2(back[i].cpu().numpy(), o.inverse(alm[i].cpu().numpy().astype(np.complex128)))
        print(f"[mixed nside={n}] forward {err_a:.2e} inverse {err_f:.2e}", flush=True)
        assert err_a < 2e-5 and err_f < 2e-5


C4_NSIDES = (64, 128, 256, 512)
C4_FIELDS = 32


@pytest.mark.timeout(1200)
def test_mixed_request_at_full_sizes():
    """Config: 32 fields each at nside 64, 128, 256 and 512, lmax = mmax =
    2 nside + 1, split-BF16, autotuned, forward then inverse in one graph
    launch.  Inputs are band-limited (a_lm with C_l = (l+1)^-2, seed
    10 + log2 nside, synthesised by the FP32 inverse).  Every size's outputs
    equal the per-size SHTCUDA / iSHTCUDA calls bit for bit, and the first and
    the last field of each size match the streamed FP64 oracle within the
    split-BF16 tolerance (forward on the FP32 input, inverse on the graph's
    own coefficients)."""
    specs = [(n, C4_FIELDS, 2 * n + 1, 2 * n + 1) for n in C4_NSIDES]
    print(f"[mixed C4] building {specs}", flush=True)
    mixed = cuhpx_amd.MixedSHT(specs, precision="bf16x3", direction="both", autotune=True)
    for i, (n, b, L, _) in enumerate(specs):
        a = band_limited_alm(L, L, b, 10 + int(np.log2(n))).astype(np.complex64)
        mixed.maps_in[i].copy_(cuhpx_amd.iSHTCUDA(n, lmax=L, mmax=L)(torch.from_numpy(a).cuda()))
    mixed.launch()
    torch.cuda.synchronize()
    alm = [a.clone() for a in mixed.alm]
    back = [g.clone() for g in mixed.maps_out]
    chk = [0, C4_FIELDS - 1]
    for i, (n, b, L, _) in enumerate(specs):
        x = mixed.maps_in[i]
        sht = cuhpx_amd.SHTCUDA(n, lmax=L, mmax=L, precision="bf16x3")
        isht = cuhpx_amd.iSHTCUDA(n, lmax=L, mmax=L, precision="bf16x3")
        assert torch.equal(alm[i], sht(x)), f"nside {n} forward differs from the per-size call"
        assert torch.equal(back[i], isht(alm[i])), f"nside {n} inverse differs from the per-size call"
        o = S.SHTOracle.healpix(n, L, L, streamed=True)
        a_ref, g_ref = o.forward_and_inverse(x[chk].double().cpu().numpy(),
                                             alm[i][chk].cpu().numpy().astype(np.complex128))
        a_got = alm[i][chk].cpu().numpy()
        g_got = back[i][chk].cpu().numpy()
        for j, fld in enumerate(chk):
            ea, ef = rel_l2(a_got[j], a_ref[j]), rel_l2(g_got[j], g_ref[j])
            print(f"[mixed C4 nside={n} field {fld}] forward {ea:.2e} inverse {ef:.2e}", flush=True)
            assert ea < 2e-5 and ef < 2e-5, (n, fld, ea, ef)


def test_mixed_graph_replays_match_eager_results():
    """Several replays with eager work and synchronisation in between; every
    replay reproduces the first."""
    mixed = cuhpx_amd.MixedSHT([(8, 3), (16, 2)], precision="fp32", direction="both")
    for t in mixed.maps_in:
        t.copy_(torch.randn_like(t))
    mixed.launch()
    torch.cuda.synchronize()
    ref = [a.clone() for a in mixed.alm] + [g.clone() for g in mixed.maps_out]
    for rep in range(4):
        for t in mixed.alm + mixed.maps_out:
            t.fill_(float("nan"))
        _ = torch.randn(1000, device="cuda").sum().item()  # eager work between replays
        mixed.launch()
        torch.cuda.synchronize()
        got = list(mixed.alm) + list(mixed.maps_out)
        for r, g in zip(ref, got):
            assert torch.equal(r, g), f"replay {rep} differs"


@pytest.mark.parametrize("direction", ["forward", "inverse"])
def test_mixed_single_direction(direction):
    specs = [(8, 2, 17, 17), (16, 3, 33, 33)]
    mixed = cuhpx_amd.MixedSHT(specs, precision="fp32", direction=direction, autotune=False)
    if direction == "forward":
        xs = [torch.randn(b, 12 * n * n, device="cuda") for n, b, _, _ in specs]
        alm, _ = mixed(xs)
        for (n, b, L, _), x, a in zip(specs, xs, alm):
            assert torch.equal(a, cuhpx_amd.SHTCUDA(n, lmax=L, mmax=L)(x))
    else:
        alms = [torch.from_numpy(band_limited_alm(L, L, b, 3).astype(np.complex64)).cuda()
                for n, b, L, _ in specs]
        _, maps = mixed(alms)
        for (n, b, L, _), a, g in zip(specs, alms, maps):
            assert torch.equal(g, cuhpx_amd.iSHTCUDA(n, lmax=L, mmax=L)(a))


def test_mixed_rejects_bad_requests():
    with pytest.raises(ValueError):
        cuhpx_amd.MixedSHT([], precision="bf16x3")
    with pytest.raises(ValueError):
        cuhpx_amd.MixedSHT([(8, 2)], precision="fp16")
    with pytest.raises(ValueError):
        cuhpx_amd.MixedSHT([(8, 0)])


def test_process_exit_with_live_graph_and_plans():
    """A process that ends with a launched mixed request and its plans still
    alive (no synchronisation, no explicit release) exits cleanly: after the
    interpreter starts to exit, the finalisers make no native calls."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import torch, cuhpx_amd\n"
            "m = cuhpx_amd.MixedSHT([(8, 2), (16, 3)], precision='bf16x3', direction='both', autotune=False)\n"
            "m.launch()\n"
            "sht = cuhpx_amd.SHTCUDA(16)\n"
            "a = sht(torch.randn(3, 12 * 16 * 16, device='cuda'))\n"
            "print('launched', flush=True)\n")
    res = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    assert "launched" in res.stdout

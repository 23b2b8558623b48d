"""The allocation-free forward SHT entry point (hpx_sht_forward_noalloc).

It runs the forward SHT with the library's fixed device scratch array as the
workspace; its results must equal hpx_sht_forward's (caller workspace) bit
for bit, on consecutive calls whose sizes shrink (the array keeps the last
call's contents), and requests larger than the array are refused before
anything runs.
"""

import numpy as np
import pytest
import torch

from cuhpx_amd import _lib
from cuhpx_amd.hpx_sht import ShtPlan

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20


@pytest.mark.parametrize("nside,L,sizes", [(256, 513, (16, 3)), (32, 65, (40, 1))])
@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_noalloc_forward_equals_workspace_forward(nside, L, sizes, precision):
    lib = _lib.lib()
    plan = ShtPlan.healpix(nside, L, L)
    s = _lib.stream_ptr()
    g = torch.Generator(device="cuda").manual_seed(nside)
    for n in sizes:
        print(f"[noalloc] nside={nside} L={L} {precision} maps={n}", flush=True)
        assert lib.hpx_sht_workspace_bytes(plan.handle, n) <= SCRATCH_BYTES
        f = torch.randn(n, plan.npix, device="cuda", generator=g)
        got = torch.full((n, L, L), float("nan"), dtype=torch.complex64, device="cuda")
        rc = lib.hpx_sht_forward_noalloc(plan.handle, _lib.PRECISIONS[precision], _lib.ptr(f), _lib.ptr(got), n, s)
        _lib.check(rc, "forward without allocation")
        want = plan.forward(f, precision)
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"maps={n}"


def test_noalloc_forward_refuses_requests_beyond_the_array():
    lib = _lib.lib()
    plan = ShtPlan.healpix(256, 513, 513)
    n = 17  # 64 GEMM columns: the workspace exceeds 128 MiB
    assert lib.hpx_sht_workspace_bytes(plan.handle, n) > SCRATCH_BYTES
    rc = lib.hpx_sht_forward_noalloc(plan.handle, 1, None, None, n, _lib.stream_ptr())
    assert rc == _lib.HPX_ERR_UNSUPPORTED
    assert "scratch" in _lib.last_error()
    assert lib.hpx_sht_forward_noalloc(None, 1, None, None, 1, _lib.stream_ptr()) == _lib.HPX_ERR_ARG
    # zero maps: nothing to do
    assert lib.hpx_sht_forward_noalloc(plan.handle, 1, None, None, 0, _lib.stream_ptr()) == _lib.HPX_OK
